"""ORACLE — test infrastructure only (tests/, __graft_entry__.smoke(), bench.py's
CPU-baseline leg). ctypes bindings of the CPU restatement (liboracle.so) and
of the reference compiled from its own sources (_ref/libyrref.so, present
only where it was built)."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REFERENCE_ROOT = os.environ.get("YAORAY_REFERENCE_ROOT", "/root/reference")

_VP, _I32, _I64 = C.c_void_p, C.c_int32, C.c_int64


def build(with_reference: bool | None = None, jobs: int = 8):
    """make -C oracle [ref dropin]; the reference parts only where its sources exist (dropin links
    yaoray_amd/lib/libyrh.so, so build that first)."""
    subprocess.run(["make", "-s", "-C", HERE, f"-j{jobs}", "all"], check=True)
    if with_reference is None:
        with_reference = os.path.isdir(os.path.join(REFERENCE_ROOT, "src"))
    if with_reference:
        subprocess.run(["make", "-s", "-C", HERE, f"-j{jobs}", "ref", "dropin", f"REFERENCE_ROOT={REFERENCE_ROOT}"],
                       check=True)


_oracle = None
_ref = None


def oracle_lib():
    global _oracle
    if _oracle is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: run oracle.oracle_py.build()")
        lib = C.CDLL(path)
        lib.oracle_trace_flat.argtypes = [_VP, _I64, _VP, _I64, _VP, _I32, _VP, _I32, _VP, _VP, _VP, _I64, _I32,
                                          _I32, _VP, _VP, _VP, _VP, _VP]
        lib.oracle_trace_two_level.argtypes = [_VP, _I64, _VP, _I64, _VP, _I32, _VP, _I32, _VP, _I32, _VP, _I32,
                                               _VP, _VP, _I32, _VP, _VP, _I64, _I32, _I32, _VP, _VP, _VP, _VP, _VP,
                                               _VP]
        lib.oracle_brute_force.argtypes = [_VP, _I64, _VP, _I64, _VP, _I32, _VP, _I32, _VP, _VP, _I64, _I32, _VP]
        _oracle = lib
    return _oracle


def ref_available() -> bool:
    return os.path.exists(os.path.join(HERE, "_ref", "libyrref.so"))


def ref_lib():
    global _ref
    if _ref is None:
        path = os.path.join(HERE, "_ref", "libyrref.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing (the reference is built only where its sources exist)")
        lib = C.CDLL(path)
        lib.ref_build.argtypes = [_VP, _I64, _VP, _I64, _VP, _I32, _VP, _VP, _I32, _I32, C.POINTER(_VP),
                                  C.POINTER(C.c_double)]
        lib.ref_free.argtypes = [_VP]
        lib.ref_free.restype = None
        lib.ref_accel_info.argtypes = [_VP] + [C.POINTER(_I32)] * 7
        lib.ref_bvh_sizes.argtypes = [_VP, _I32] + [C.POINTER(_I32)] * 4
        lib.ref_bvh_copy.argtypes = [_VP, _I32, _VP, _VP, _VP]
        lib.ref_tlas_copy.argtypes = [_VP, _VP, _VP, _VP, _VP]
        lib.ref_trace.argtypes = [_VP, _VP, _VP, _I64, _I32, _VP, _VP, _VP, _VP, _VP]
        lib.ref_inverse.argtypes = [_VP, _VP]
        lib.ref_inverse.restype = None
        _ref = lib
    return _ref


def _p(a):
    if a is None or len(a) == 0:
        return None
    assert a.flags["C_CONTIGUOUS"], "oracle arrays must be C-contiguous"
    return a.ctypes.data


def _c(a, dtype=np.float32):
    return np.ascontiguousarray(a, dtype=dtype)


# -- restated oracle ------------------------------------------------------------

def _spheres(spheres):
    if spheres is None or len(spheres) == 0:
        return None, 0
    assert spheres.dtype.itemsize == 28 and spheres.flags["C_CONTIGUOUS"]
    return spheres.ctypes.data, len(spheres)


def trace_flat(vertices, indices, primitives, bvh, org, dirs, occlusion=False, threads=8, spheres=None):
    """Returns dict(t, u, v, triangle, node_tests, triangle_tests, sphere_tests, occluded);
    `triangle` is the hit id (-1 miss, >= 0 flat triangle, -2 - s sphere s)."""
    org, dirs = _c(org), _c(dirs)
    n = len(org)
    hit = np.zeros((n, 4), np.float32)
    occ = np.zeros(n, np.uint32)
    nt = np.zeros(n, np.uint32)
    tt = np.zeros(n, np.uint32)
    st = np.zeros(n, np.uint32)
    desc = bvh.desc()
    sp, ns = _spheres(spheres)
    oracle_lib().oracle_trace_flat(_p(vertices), len(vertices), _p(indices), len(indices), _p(primitives),
                                   len(primitives), sp, ns, C.cast(C.pointer(desc), C.c_void_p),
                                   org.ctypes.data, dirs.ctypes.data, n, int(occlusion), threads,
                                   hit.ctypes.data, occ.ctypes.data, nt.ctypes.data, tt.ctypes.data, st.ctypes.data)
    return {"t": hit[:, 0].copy(), "u": hit[:, 1].copy(), "v": hit[:, 2].copy(),
            "triangle": hit[:, 3].copy().view(np.int32), "node_tests": nt, "triangle_tests": tt,
            "sphere_tests": st, "occluded": occ}


def trace_two_level(vertices, indices, primitives, blases, instances, tlas, tlas_prims, org, dirs,
                    occlusion=False, threads=8, spheres=None):
    from yaoray_amd import native as N
    org, dirs = _c(org), _c(dirs)
    n = len(org)
    hit = np.zeros((n, 4), np.float32)
    inst = np.zeros(n, np.int32)
    occ = np.zeros(n, np.uint32)
    nt = np.zeros(n, np.uint32)
    tt = np.zeros(n, np.uint32)
    st = np.zeros(n, np.uint32)
    descs = (N.BvhDesc * max(1, len(blases)))(*[b.desc() for b in blases])
    tdesc = tlas.desc()
    sp, ns = _spheres(spheres)
    oracle_lib().oracle_trace_two_level(_p(vertices), len(vertices), _p(indices), len(indices), _p(primitives),
                                        len(primitives), sp, ns, C.cast(descs, C.c_void_p), len(blases),
                                        _p(instances), len(instances), C.cast(C.pointer(tdesc), C.c_void_p),
                                        _p(tlas_prims), len(tlas_prims), org.ctypes.data, dirs.ctypes.data, n,
                                        int(occlusion), threads, hit.ctypes.data, inst.ctypes.data,
                                        occ.ctypes.data, nt.ctypes.data, tt.ctypes.data, st.ctypes.data)
    return {"t": hit[:, 0].copy(), "u": hit[:, 1].copy(), "v": hit[:, 2].copy(),
            "triangle": hit[:, 3].copy().view(np.int32), "instance": inst, "node_tests": nt,
            "triangle_tests": tt, "sphere_tests": st, "occluded": occ}


def brute_force(vertices, indices, primitives, org, dirs, threads=8, spheres=None):
    org, dirs = _c(org), _c(dirs)
    n = len(org)
    hit = np.zeros((n, 4), np.float32)
    sp, ns = _spheres(spheres)
    oracle_lib().oracle_brute_force(_p(vertices), len(vertices), _p(indices), len(indices), _p(primitives),
                                    len(primitives), sp, ns, org.ctypes.data, dirs.ctypes.data, n, threads,
                                    hit.ctypes.data)
    return {"t": hit[:, 0].copy(), "u": hit[:, 1].copy(), "v": hit[:, 2].copy(),
            "triangle": hit[:, 3].copy().view(np.int32)}


# -- the reference itself -----------------------------------------------------------

def sample_analytic(scene, points, u):
    """The oracle's SampleAnalyticLight on the scene's analytic lights and
    tables: dict(wi, distance, radiance, selection_pdf, light_index, valid)."""
    lib = oracle_lib()
    lib.oracle_sample_analytic.argtypes = [_VP, _I64, _VP, _VP, _VP]
    p = _c(np.asarray(points).reshape(-1, 3))
    uu = _c(np.asarray(u).reshape(-1))
    out = np.zeros((len(p), 10), np.float32)
    sd = scene.shading_desc()
    if lib.oracle_sample_analytic(C.cast(C.pointer(sd), _VP), len(p), p.ctypes.data, uu.ctypes.data,
                                  out.ctypes.data) != 0:
        raise RuntimeError("oracle_sample_analytic failed")
    return {"wi": out[:, 0:3], "distance": out[:, 3], "radiance": out[:, 4:7], "selection_pdf": out[:, 7],
            "light_index": out[:, 8].astype(np.int32), "valid": out[:, 9] > 0}


def _ref_analytic(lights):
    """lights: ctypes array of native.AnalyticLight (or None)."""
    lib = ref_lib()
    lib.ref_set_analytic_lights.argtypes = [_VP, _I32]
    lib.ref_set_analytic_lights.restype = None
    if lights is None or len(lights) == 0:
        lib.ref_set_analytic_lights(None, 0)
    else:
        lib.ref_set_analytic_lights(C.cast(lights, _VP), len(lights))


def _scene_analytic(scene):
    from yaoray_amd import native as N
    sd = scene.shading_desc()
    n = sd.analytic_light_count
    if n == 0:
        return None
    return (N.AnalyticLight * n).from_buffer_copy(C.string_at(sd.analytic_lights, n * C.sizeof(N.AnalyticLight)))


def ref_analytic_tables(lights):
    """The reference's PrepareLightSampling tables for `lights` (ctypes array of
    native.AnalyticLight): (probability, alias, pmf, tree nodes (LIGHT_NODE_DTYPE),
    root, use_tree)."""
    lib = ref_lib()
    lib.ref_analytic_tables.argtypes = [_VP, _I32, _VP, _VP, _VP, _VP, C.POINTER(_I32), C.POINTER(_I32)]
    lib.ref_analytic_tables.restype = _I32
    n = len(lights)
    prob, alias, pmf = np.zeros(n, np.float32), np.zeros(n, np.int32), np.zeros(n, np.float32)
    tree = np.zeros(max(1, 2 * n), LIGHT_NODE_DTYPE)
    root, use = _I32(), _I32()
    count = lib.ref_analytic_tables(C.cast(lights, _VP), n, prob.ctypes.data, alias.ctypes.data, pmf.ctypes.data,
                                    tree.ctypes.data, C.byref(root), C.byref(use))
    return prob, alias, pmf, tree[:count].copy(), root.value, use.value


LIGHT_NODE_DTYPE = np.dtype([("bounds_min", "<f4", 3), ("bounds_max", "<f4", 3), ("power", "<f4"),
                             ("left", "<i4"), ("right", "<i4"), ("light_index", "<i4")])


def _ref_spheres(spheres):
    lib = ref_lib()
    lib.ref_set_spheres.argtypes = [_VP, _I32]
    lib.ref_set_spheres.restype = None
    sp, n = _spheres(spheres)
    lib.ref_set_spheres(sp, n)


class RefScene:
    """The reference's RenderSceneIR + RenderAcceleration built by its own code."""

    def __init__(self, vertices, indices, primitives, instance_primitive=None, instance_matrices=None,
                 threads=0, spheres=None):
        lib = ref_lib()
        self._lib = lib
        self._h = None
        _ref_spheres(spheres)
        ninst = 0 if instance_primitive is None else len(instance_primitive)
        ip = np.ascontiguousarray(instance_primitive if ninst else np.zeros(1), np.int32)
        im = np.ascontiguousarray(instance_matrices if ninst else np.zeros((1, 16)), np.float32)
        h = C.c_void_p()
        secs = C.c_double()
        rc = lib.ref_build(_p(vertices), len(vertices), _p(indices), len(indices), _p(primitives), len(primitives),
                           ip.ctypes.data, im.ctypes.data, ninst, threads, C.byref(h), C.byref(secs))
        _ref_spheres(None)
        if rc != 0:
            raise RuntimeError("reference BuildRenderAcceleration reported errors")
        self._h = h
        self.build_seconds = secs.value

    def __del__(self):
        try:
            if self._h:
                self._lib.ref_free(self._h)
        except Exception:
            pass

    def info(self):
        vals = [C.c_int32() for _ in range(7)]
        self._lib.ref_accel_info(self._h, *[C.byref(v) for v in vals])
        keys = ["kind", "nodes", "prim_indices", "prims", "max_depth", "blas_count", "instance_count"]
        return {k: v.value for k, v in zip(keys, vals)}

    def bvh(self, blas=-1):
        from yaoray_amd.api import NODE_DTYPE, PRIM_DTYPE, BvhArrays
        n, pi, pr, md = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
        self._lib.ref_bvh_sizes(self._h, blas, C.byref(n), C.byref(pi), C.byref(pr), C.byref(md))
        nodes = np.zeros(n.value, NODE_DTYPE)
        pidx = np.zeros(pi.value, np.int32)
        prims = np.zeros(pr.value, PRIM_DTYPE)
        self._lib.ref_bvh_copy(self._h, blas, _p(nodes) or 0, _p(pidx) or 0, _p(prims) or 0)
        return BvhArrays(nodes, pidx, prims, md.value)

    def tlas(self):
        from yaoray_amd.api import NODE_DTYPE
        info = self.info()
        nodes = np.zeros(info["nodes"], NODE_DTYPE)
        pidx = np.zeros(info["prim_indices"], np.int32)
        prim_inst = np.zeros(info["prims"], np.int32)
        inst = np.zeros((info["instance_count"], 40), np.float32)
        self._lib.ref_tlas_copy(self._h, _p(nodes) or 0, _p(pidx) or 0, _p(prim_inst) or 0, _p(inst) or 0)
        return nodes, pidx, prim_inst, inst

    def trace(self, org, dirs, threads=8):
        org, dirs = _c(org), _c(dirs)
        n = len(org)
        hit = np.zeros((n, 4), np.float32)
        inst = np.zeros(n, np.int32)
        nt = np.zeros(n, np.uint32)
        tt = np.zeros(n, np.uint32)
        st = np.zeros(n, np.uint32)
        self._lib.ref_trace(self._h, org.ctypes.data, dirs.ctypes.data, n, threads, hit.ctypes.data,
                            inst.ctypes.data, nt.ctypes.data, tt.ctypes.data, st.ctypes.data)
        return {"t": hit[:, 0].copy(), "u": hit[:, 1].copy(), "v": hit[:, 2].copy(),
                "triangle": hit[:, 3].copy().view(np.int32), "instance": inst, "node_tests": nt,
                "triangle_tests": tt, "sphere_tests": st}


STAT_KEYS = ("rays_traced", "shadow_rays", "occluded_shadow_rays", "triangle_tests", "sphere_tests",
             "bvh_node_tests", "hits", "misses", "bad_samples")


def render(scene, width, height, spp, max_depth=8, sampler=2, seed=0x59414F524159, radiance_clamp=0.0,
           threads=8, first_sample=0):
    """The restated CPU path tracer on the scene's own C-ABI descriptors.
    Returns (FP32 film records, stats dict)."""
    from yaoray_amd.api import FILM32_DTYPE

    lib = oracle_lib()
    if not hasattr(lib, "_render_sig"):
        lib.oracle_render.argtypes = [_VP, _VP, _VP, _I32, _I32, _I32, _VP, _VP]
        lib.oracle_sample_values.argtypes = [_I32, C.c_uint64, _I32, _I64, _VP, _VP, _VP, _VP, _I32, _VP]
        lib._render_sig = True
    gd = scene.geometry_desc()
    sd = scene.shading_desc()
    rd = scene.render_desc(width, height, spp, max_depth, sampler, seed, radiance_clamp, 0)
    film = np.zeros(width * height, FILM32_DTYPE)
    stats = np.zeros(9, np.uint64)
    rc = lib.oracle_render(C.cast(C.pointer(gd), _VP), C.cast(C.pointer(sd), _VP), C.cast(C.pointer(rd), _VP),
                           first_sample, spp - first_sample, threads, film.ctypes.data, stats.ctypes.data)
    if rc != 0:
        raise RuntimeError("oracle_render failed")
    return film, dict(zip(STAT_KEYS, (int(x) for x in stats)))


def render_adaptive(scene, width, height, spp, adaptive, max_depth=8, sampler=2, seed=0x59414F524159,
                    radiance_clamp=0.0, threads=8, film=None, first_pass=0):
    """The restated pass loop with active pixels (adaptive sampling when
    adaptive = (min_spp, relative_error, absolute_error, confidence); None turns it
    off). Continues `film` (FP32 records) from pass `first_pass` when given.
    Returns (film, stats dict, passes run, converged pixels)."""
    from yaoray_amd.api import FILM32_DTYPE

    lib = oracle_lib()
    lib.oracle_render_adaptive.argtypes = [_VP, _VP, _VP, _VP, _I32, _I32, _VP, _VP, C.POINTER(_I32),
                                           C.POINTER(C.c_uint64)]
    gd = scene.geometry_desc()
    sd = scene.shading_desc()
    rd = scene.render_desc(width, height, spp, max_depth, sampler, seed, radiance_clamp, 0)
    film = np.zeros(width * height, FILM32_DTYPE) if film is None else np.array(film, FILM32_DTYPE)
    a = np.array([0.0, 0.0, 0.0, 0.0, 0.0] if adaptive is None else [1.0, *adaptive], np.float32)
    stats = np.zeros(9, np.uint64)
    passes, conv = _I32(), C.c_uint64()
    rc = lib.oracle_render_adaptive(C.cast(C.pointer(gd), _VP), C.cast(C.pointer(sd), _VP), C.cast(C.pointer(rd), _VP),
                                    a.ctypes.data, first_pass, threads, film.ctypes.data, stats.ctypes.data,
                                    C.byref(passes), C.byref(conv))
    if rc != 0:
        raise RuntimeError("oracle_render_adaptive failed")
    return film, dict(zip(STAT_KEYS, (int(x) for x in stats))), passes.value, conv.value


def sample_values(kind, seed, spp, x, y, sample, dim, dims):
    lib = oracle_lib()
    lib.oracle_sample_values.argtypes = [_I32, C.c_uint64, _I32, _I64, _VP, _VP, _VP, _VP, _I32, _VP]
    arrs = [np.ascontiguousarray(a, np.int32) for a in (x, y, sample, dim)]
    out = np.zeros(len(x) * dims, np.float32)
    lib.oracle_sample_values(kind, seed, spp, len(x), *[a.ctypes.data for a in arrs], dims, out.ctypes.data)
    return out.reshape(len(x), dims) if dims == 2 else out


def ref_render(scene, width, height, spp, max_depth=8, sampler=2, seed=0x59414F524159, radiance_clamp=0.0,
               threads=8, adaptive=None):
    """The reference's own PrepareCpuScene + RenderCpuPathTrace on the same inputs, with
    adaptive = (min_spp, relative_error, absolute_error, confidence) or None (off).
    Returns (FP32 film records, stats dict with converged_pixels, render seconds)."""
    from yaoray_amd.api import FILM32_DTYPE

    lib = ref_lib()
    lib.ref_set_adaptive.argtypes = [_I32, _I32, C.c_float, C.c_float, C.c_float]
    lib.ref_last_converged.restype = C.c_uint64
    if adaptive is None:
        lib.ref_set_adaptive(0, 16, 0.05, 0.001, 1.96)
    else:
        lib.ref_set_adaptive(1, int(adaptive[0]), *[float(v) for v in adaptive[1:]])
    lib.ref_render.argtypes = [_VP, _I64, _VP, _I64, _VP, _I32, _VP, _VP, _I32, _VP, _VP, _VP, _I32, _VP, _VP,
                               C.POINTER(C.c_double)]
    v, i, p = scene.geometry()
    info = scene.info
    if info.instance_count:
        _, _, inst = scene.tlas()
        ip = np.ascontiguousarray(inst["primitive"], np.int32)
        im = np.ascontiguousarray(inst["object_to_world"], np.float32)
    else:
        ip = np.zeros(1, np.int32)
        im = np.zeros((1, 16), np.float32)
    sd = scene.shading_desc()
    rd = scene.render_desc(width, height, spp, max_depth, sampler, seed, radiance_clamp, 0)
    cam = np.array(list(info.camera), np.float32)
    film = np.zeros(width * height, FILM32_DTYPE)
    stats = np.zeros(10, np.uint64)
    secs = C.c_double()
    _ref_spheres(scene.spheres())
    _ref_analytic(_scene_analytic(scene))
    rc = lib.ref_render(_p(v), len(v), _p(i), len(i), _p(p), len(p), ip.ctypes.data, im.ctypes.data,
                        info.instance_count, C.cast(C.pointer(sd), _VP), cam.ctypes.data, C.cast(C.pointer(rd), _VP),
                        threads, film.ctypes.data, stats.ctypes.data, C.byref(secs))
    _ref_spheres(None)
    _ref_analytic(None)
    if rc != 0:
        raise RuntimeError(f"reference render failed ({rc})")
    out = dict(zip(STAT_KEYS, (int(x) for x in stats[:9])))
    out["threads"] = int(stats[9])
    out["converged_pixels"] = int(lib.ref_last_converged())
    lib.ref_set_adaptive(0, 16, 0.05, 0.001, 1.96)
    return film, out, secs.value


def _texture(texels, filter, wrap_s, wrap_t):
    from yaoray_amd import native as N

    t = np.ascontiguousarray(texels, np.float32)
    d = N.Texture()
    d.height, d.width = (int(t.shape[0]), int(t.shape[1])) if t.size else (0, 0)
    d.filter, d.wrap_s, d.wrap_t = int(filter), int(wrap_s), int(wrap_s if wrap_t is None else wrap_t)
    d.texels = t.ctypes.data if t.size else None
    return d, t


def _sample_texture(lib, name, texels, uv, footprint, filter, wrap_s, wrap_t):
    d, keep = _texture(texels, filter, wrap_s, wrap_t)
    uv = _c(uv)
    fp = None if footprint is None else _c(footprint)
    out = np.zeros((len(uv), 4), np.float32)
    fn = getattr(lib, name)
    fn.argtypes = [_VP, _I64, _VP, _VP, _VP]
    fn(C.cast(C.pointer(d), _VP), len(uv), _p(uv), None if fp is None else _p(fp), _p(out))
    return out


def sample_texture(texels, uv, footprint=None, filter=1, wrap_s=0, wrap_t=None):
    """The oracle's SampleTexture4 of an RGBA texture (height x width x 4) at uv
    (n x 2) with footprints (n x 4: dudx, dvdx, dudy, dvdy) or none; n x 4."""
    return _sample_texture(oracle_lib(), "oracle_sample_texture", texels, uv, footprint, filter, wrap_s, wrap_t)


def ref_sample_texture(texels, uv, footprint=None, filter=1, wrap_s=0, wrap_t=None):
    """The reference's SampleTexture4 (after BuildTextureSamplingCaches) on the same inputs."""
    return _sample_texture(ref_lib(), "ref_sample_texture", texels, uv, footprint, filter, wrap_s, wrap_t)


def _mips(lib, name, texels, wrap_s, wrap_t):
    d, keep = _texture(texels, 1, wrap_s, wrap_t)
    fn = getattr(lib, name)
    fn.argtypes = [_VP, C.POINTER(_I32), C.POINTER(_I64), _VP]
    levels, count = _I32(), _I64()
    fn(C.cast(C.pointer(d), _VP), C.byref(levels), C.byref(count), None)
    out = np.zeros((count.value, 4), np.float32)
    fn(C.cast(C.pointer(d), _VP), C.byref(levels), C.byref(count), _p(out))
    return levels.value, out


def texture_mips(texels, wrap_s=0, wrap_t=None):
    """(level count, every level's texels one after another) of the oracle's mip chain."""
    return _mips(oracle_lib(), "oracle_texture_mips", texels, wrap_s, wrap_t)


def ref_texture_mips(texels, wrap_s=0, wrap_t=None):
    """The same from the reference's BuildTextureMipChain."""
    return _mips(ref_lib(), "ref_texture_mips", texels, wrap_s, wrap_t)


def ref_sample_values(kind, seed, spp, x, y, sample, dim, dims):
    """The reference's yr::Sampler at bounce 0 (Sample1D / Sample2D of raw dimensions)."""
    lib = ref_lib()
    lib.ref_sample_values.argtypes = [_I32, C.c_uint64, _I32, _I64, _VP, _VP, _VP, _VP, _I32, _VP]
    arrs = [np.ascontiguousarray(a, np.int32) for a in (x, y, sample, dim)]
    out = np.zeros(len(x) * dims, np.float32)
    lib.ref_sample_values(kind, seed, spp, len(x), *[a.ctypes.data for a in arrs], dims, out.ctypes.data)
    return out.reshape(len(x), dims) if dims == 2 else out


def ref_inverse(m):
    m = np.ascontiguousarray(m, np.float32).reshape(16)
    out = np.zeros(16, np.float32)
    ref_lib().ref_inverse(m.ctypes.data, out.ctypes.data)
    return out


def camera_rays(render_desc):
    lib = oracle_lib()
    lib.oracle_camera_rays.argtypes = [_VP, _VP, _VP]
    n = render_desc.width * render_desc.height
    org = np.zeros((n, 4), np.float32)
    dirs = np.zeros((n, 4), np.float32)
    lib.oracle_camera_rays(C.cast(C.pointer(render_desc), _VP), org.ctypes.data, dirs.ctypes.data)
    return org, dirs


# ---- learned guiding (oracle_guide.cpp) ------------------------------------------------


def _guide_lib():
    lib = oracle_lib()
    if not hasattr(lib, "_guide_sig"):
        lib.oracle_guide_config.argtypes = [_VP, _VP, _VP]
        lib.oracle_guide_forward.argtypes = [_VP, _VP, _VP, _I64, _VP, _I32, _VP]
        lib.oracle_guide_gradients.argtypes = [_VP, _VP, _VP, _I64, _VP, C.c_double, _I32, _VP, _VP, _VP]
        lib.oracle_guide_loss.argtypes = [_VP, _VP, _VP, _I64, _VP, C.c_double]
        lib.oracle_guide_loss.restype = C.c_double
        lib.oracle_guide_mixture_eval.argtypes = [_I64, _VP, _VP, _VP, _VP, _VP]
        lib._guide_sig = True
    return lib


def guide_config(desc):
    res, dense = np.zeros(8, np.int32), np.zeros(8, np.int32)
    _guide_lib().oracle_guide_config(C.cast(C.pointer(desc), _VP), res.ctypes.data, dense.ctypes.data)
    return res, dense


def guide_forward(desc, mlp, grid, rows, bf16=False):
    r = _c(np.asarray(rows, np.float32).reshape(-1, 9))
    m, g = _c(np.asarray(mlp, np.float32)), _c(np.asarray(grid, np.float32))
    out = np.zeros((len(r), 40), np.float32)
    _guide_lib().oracle_guide_forward(C.cast(C.pointer(desc), _VP), m.ctypes.data, g.ctypes.data, len(r),
                                      r.ctypes.data, 1 if bf16 else 0, out.ctypes.data)
    return out


def guide_gradients(desc, mlp, grid, samples, w_norm, bf16=False):
    s = _c(np.asarray(samples, np.float32).reshape(-1, 13))
    m, g = _c(np.asarray(mlp, np.float32)), _c(np.asarray(grid, np.float32))
    gm, gg, loss = np.zeros(len(m)), np.zeros(len(g)), C.c_double()
    _guide_lib().oracle_guide_gradients(C.cast(C.pointer(desc), _VP), m.ctypes.data, g.ctypes.data, len(s),
                                        s.ctypes.data, w_norm, 1 if bf16 else 0, gm.ctypes.data, gg.ctypes.data,
                                        C.byref(loss))
    return gm, gg, loss.value


def guide_loss(desc, mlp, grid, samples, w_norm):
    s = _c(np.asarray(samples, np.float32).reshape(-1, 13))
    m, g = _c(np.asarray(mlp, np.float32)), _c(np.asarray(grid, np.float32))
    return _guide_lib().oracle_guide_loss(C.cast(C.pointer(desc), _VP), m.ctypes.data, g.ctypes.data, len(s),
                                          s.ctypes.data, w_norm)


def guide_mixture_eval(mixture, dirs, u):
    m = _c(np.asarray(mixture, np.float32).reshape(-1, 40))
    d = _c(np.asarray(dirs, np.float32).reshape(-1, 3))
    uu = _c(np.asarray(u, np.float32).reshape(-1, 3))
    pdf, smp = np.zeros(len(m), np.float32), np.zeros((len(m), 4), np.float32)
    _guide_lib().oracle_guide_mixture_eval(len(m), m.ctypes.data, d.ctypes.data, uu.ctypes.data, pdf.ctypes.data,
                                           smp.ctypes.data)
    return pdf, smp


# ---- BSDF queries (query layout of oracle_bsdf_query) ---------------------------------


def _bsdf(fn_name, lib, materials, material_index, queries):
    from yaoray_amd.api import material_array
    arr = material_array(materials)
    idx = np.ascontiguousarray(material_index, np.int32)
    q = np.ascontiguousarray(queries, np.float32).reshape(-1, 12)
    out = np.zeros((len(q), 12), np.float32)
    fn = getattr(lib, fn_name)
    fn.argtypes = [_VP, _I32, _I64, _VP, _VP, _VP]
    if fn(C.cast(arr, _VP), len(materials), len(q), idx.ctypes.data, q.ctypes.data, out.ctypes.data) != 0:
        raise RuntimeError(f"{fn_name} failed")
    return out


def bsdf_query(materials, material_index, queries):
    """The restated EvaluateBsdf / PdfBsdf / SampleBsdf / IsDeltaBsdf per query row
    (wo, wi, normal, u0, u1, unused) -> (f, pdf, sampled wi, weight, pdf, flags)."""
    return _bsdf("oracle_bsdf_query", oracle_lib(), materials, material_index, queries)


def ref_bsdf_query(materials, material_index, queries):
    """The reference's own BSDF functions on the same query rows."""
    return _bsdf("ref_bsdf_query", ref_lib(), materials, material_index, queries)
