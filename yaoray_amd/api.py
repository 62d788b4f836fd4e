"""Python face of the backend: scenes (host library) and a device context (libyrh).

Host arrays are numpy; device arrays are ``DeviceBuffer`` objects owned by a
``Device`` and allocated through the C ABI, so this module needs neither HIP
headers nor torch.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import native as N

NODE_DTYPE = np.dtype([("bounds_min", "<f4", 3), ("bounds_max", "<f4", 3),
                       ("payload_offset", "<i4"), ("primitive_count", "<i4")])
PRIM_DTYPE = np.dtype([("kind", "<i4"), ("mesh_primitive", "<i4"), ("sphere", "<i4"),
                       ("local_triangle", "<i4"), ("flat_triangle_index", "<i4")])
VERTEX_DTYPE = np.dtype([("position", "<f4", 3), ("normal", "<f4", 3), ("uv", "<f4", 2),
                         ("tangent", "<f4", 3), ("tangent_handedness", "<f4")])
PRIMITIVE_DTYPE = np.dtype([("first_index", "<u4"), ("index_count", "<u4"), ("material_index", "<i4"),
                            ("has_normals", "u1"), ("has_uvs", "u1"), ("has_tangents", "u1"),
                            ("reserved", "u1")])
INSTANCE_DTYPE = np.dtype([("primitive", "<i4"), ("blas_index", "<i4"), ("object_to_world", "<f4", 16),
                           ("world_to_object", "<f4", 16), ("world_bounds_min", "<f4", 3),
                           ("world_bounds_max", "<f4", 3)])
TLAS_PRIM_DTYPE = np.dtype([("kind", "<i4"), ("instance", "<i4"), ("sphere", "<i4"),
                            ("world_bounds_min", "<f4", 3), ("world_bounds_max", "<f4", 3)])

# yrh_film_pixel (FP64 sums and Welford moments) and yr::FilmPixel (FP32, 56 bytes).
SPHERE_DTYPE = np.dtype([("center", "<f4", 3), ("radius", "<f4"), ("material_index", "<i4"),
                         ("area_light_index", "<i4"), ("flip_normals", "u1"), ("reserved", "u1", 3)])
FILM64_DTYPE = np.dtype([("sum", "<f8", 3), ("luminance_mean", "<f8"), ("luminance_m2", "<f8"),
                         ("albedo_sum", "<f8", 3), ("normal_sum", "<f8", 3), ("depth_sum", "<f8"),
                         ("samples", "<u4"), ("aov_samples", "<u4")])
FILM32_DTYPE = np.dtype([("sum", "<f4", 3), ("samples", "<u4"), ("luminance_mean", "<f4"), ("luminance_m2", "<f4"),
                         ("albedo_sum", "<f4", 3), ("normal_sum", "<f4", 3), ("depth_sum", "<f4"),
                         ("aov_samples", "<u4")])
assert FILM64_DTYPE.itemsize == 104 and FILM32_DTYPE.itemsize == 56

assert NODE_DTYPE.itemsize == 32 and PRIM_DTYPE.itemsize == 20 and VERTEX_DTYPE.itemsize == 48
assert PRIMITIVE_DTYPE.itemsize == 16 and INSTANCE_DTYPE.itemsize == 160 and TLAS_PRIM_DTYPE.itemsize == 36


class YrhError(RuntimeError):
    pass


def _view(ptr, dtype, count):
    """Copy `count` records of `dtype` from a C pointer into a numpy array."""
    dtype = np.dtype(dtype)
    if count == 0 or not ptr:
        return np.zeros(0, dtype=dtype)
    buf = (C.c_char * (dtype.itemsize * count)).from_address(ptr)
    return np.frombuffer(buf, dtype=dtype, count=count).copy()


def _ptr(a: np.ndarray) -> int:
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


class BvhArrays:
    """One binary BVH as the reference lays it out (yr::RenderBvh)."""

    def __init__(self, nodes, primitive_indices, primitives, max_depth):
        self.nodes = nodes
        self.primitive_indices = primitive_indices
        self.primitives = primitives
        self.max_depth = int(max_depth)

    @staticmethod
    def from_desc(d: N.BvhDesc) -> "BvhArrays":
        return BvhArrays(_view(d.nodes, NODE_DTYPE, d.node_count),
                         _view(d.primitive_indices, "<i4", d.primitive_index_count),
                         _view(d.primitives, PRIM_DTYPE, d.primitive_count), d.max_depth)

    def desc(self) -> N.BvhDesc:
        d = N.BvhDesc()
        d.nodes = _ptr(self.nodes) if len(self.nodes) else None
        d.node_count = len(self.nodes)
        d.primitive_indices = _ptr(self.primitive_indices) if len(self.primitive_indices) else None
        d.primitive_index_count = len(self.primitive_indices)
        d.primitives = _ptr(self.primitives) if len(self.primitives) else None
        d.primitive_count = len(self.primitives)
        d.max_depth = self.max_depth
        return d


def validate_geometry(desc: N.GeometryDesc):
    """yrh_validate_geometry: the device-free checks of a geometry upload.
    Returns (status, message); status 0 when the geometry would be accepted."""
    buf = C.create_string_buffer(512)
    rc = N.yrh().yrh_validate_geometry(C.byref(desc), buf, len(buf))
    return rc, buf.value.decode()


def material_array(materials):
    """ctypes array of native.Material from dicts (kind, reflectance, ior, uroughness, vroughness,
    alpha, emission, absorption_distance, absorption_color, reflectance_texture, alpha_texture,
    normal_map, normal_scale; missing keys take the defaults, -1 = no texture)."""
    arr = (N.Material * max(1, len(materials)))()
    for k, m in enumerate(materials):
        a = arr[k]
        a.kind = m.get("kind", 0)
        a.reflectance[:] = m.get("reflectance", (0.5, 0.5, 0.5))
        a.ior = m.get("ior", 1.5)
        a.uroughness = m.get("uroughness", 0.0)
        a.vroughness = m.get("vroughness", m.get("uroughness", 0.0))
        a.alpha = m.get("alpha", 1.0)
        a.emission[:] = m.get("emission", (0.0, 0.0, 0.0))
        a.absorption_distance = m.get("absorption_distance", 1.0)
        a.absorption_color[:] = m.get("absorption_color", (1.0, 1.0, 1.0))
        a.reflectance_texture = m.get("reflectance_texture", -1)
        a.alpha_texture = m.get("alpha_texture", -1)
        a.normal_map = m.get("normal_map", -1)
        a.normal_scale = m.get("normal_scale", 1.0)
    return arr


class Scene:
    """A RenderSceneIR owned by the host library, with its acceleration structure."""

    def __init__(self, handle):
        self._lib = N.host()
        self._h = handle

    # -- construction ---------------------------------------------------
    @classmethod
    def _create(cls, kind, params, seed):
        lib = N.host()
        arr = (C.c_int64 * max(1, len(params)))(*params)
        h = C.c_void_p()
        rc = lib.yrhost_scene_create(kind, arr, len(params), seed, C.byref(h))
        if rc != 0:
            raise YrhError(f"yrhost_scene_create({kind}) failed: {rc}")
        return cls(h)

    @classmethod
    def cornell(cls, material_seed: int = 1) -> "Scene":
        return cls._create(N.SCENE_CORNELL, [], material_seed)

    @classmethod
    def triangles(cls, count: int, seed: int = 2) -> "Scene":
        return cls._create(N.SCENE_TRIANGLES, [count], seed)

    @classmethod
    def instanced(cls, instances=10000, subdivisions=4, variants=8, env_width=2048, env_height=1024,
                  seed: int = 4) -> "Scene":
        return cls._create(N.SCENE_INSTANCED, [instances, subdivisions, variants, env_width, env_height], seed)

    @classmethod
    def room(cls, furniture_triangles=200000, env_width=1024, env_height=512, seed: int = 6) -> "Scene":
        """Guiding workload: window-lit room with a hidden ceiling lamp (flat BVH)."""
        return cls._create(N.SCENE_ROOM, [furniture_triangles, env_width, env_height], seed)

    @classmethod
    def from_arrays(cls, vertices, indices, primitives, instance_primitive=None, instance_matrices=None):
        lib = N.host()
        vertices = np.ascontiguousarray(vertices, dtype=VERTEX_DTYPE)
        indices = np.ascontiguousarray(indices, dtype="<u4")
        primitives = np.ascontiguousarray(primitives, dtype=PRIMITIVE_DTYPE)
        ninst = 0 if instance_primitive is None else len(instance_primitive)
        ip = np.ascontiguousarray(instance_primitive if ninst else np.zeros(1), dtype="<i4")
        im = np.ascontiguousarray(instance_matrices if ninst else np.zeros((1, 16)), dtype="<f4")
        h = C.c_void_p()
        rc = lib.yrhost_scene_from_arrays(_ptr(vertices) if len(vertices) else None, len(vertices),
                                          _ptr(indices) if len(indices) else None, len(indices),
                                          _ptr(primitives) if len(primitives) else None, len(primitives),
                                          _ptr(ip), _ptr(im), ninst, C.byref(h))
        if rc != 0:
            raise YrhError(f"yrhost_scene_from_arrays failed: {rc}")
        return cls(h)

    @classmethod
    def from_ir(cls, sections):
        """A Scene from the sections of a compiled RenderSceneIR as raw byte arrays
        (`oracle/dropin_probe.cpp dump`: vertices, indices, primitives, spheres,
        inst_prim, inst_mat, camera, materials, emissives, analytic, environment,
        settings; yrh.h records). Returns (scene with its acceleration built,
        settings dict: width, height, spp, max_depth, sampler, seed)."""
        g = lambda k: np.ascontiguousarray(sections[k], np.uint8)
        ip = g("inst_prim").view(np.int32)
        im = g("inst_mat").view(np.float32).reshape(-1, 16)
        s = cls.from_arrays(g("vertices").view(VERTEX_DTYPE), g("indices").view(np.uint32),
                            g("primitives").view(PRIMITIVE_DTYPE), ip if len(ip) else None, im if len(ip) else None)
        cam = g("camera").view(np.float32)
        s.set_camera(cam[0:3], cam[3:6], cam[6:9], cam[9:12], float(cam[12]))
        raw = g("materials").tobytes()
        mats = (N.Material * (len(raw) // C.sizeof(N.Material))).from_buffer_copy(raw)
        s.set_materials([{k: (list(getattr(m, k)) if isinstance(getattr(m, k), C.Array) else getattr(m, k))
                          for k, _ in N.Material._fields_ if k != "reserved"} for m in mats])
        sph = g("spheres").view(SPHERE_DTYPE)
        s.set_spheres([(tuple(x["center"]), float(x["radius"]), int(x["material_index"]), bool(x["flip_normals"]))
                       for x in sph])
        raw = g("emissives").tobytes()
        em = (N.Emissive * (len(raw) // C.sizeof(N.Emissive))).from_buffer_copy(raw)
        s.set_emissives([(e.primitive_index, tuple(e.radiance), e.area) for e in em])
        raw = g("analytic").tobytes()
        al = (N.AnalyticLight * (len(raw) // C.sizeof(N.AnalyticLight))).from_buffer_copy(raw)
        s.set_analytic_lights([{"kind": a.kind, "position": tuple(a.position), "direction": tuple(a.direction),
                                "intensity": tuple(a.intensity), "cone_angle": a.cone_angle,
                                "cone_cos_inner": a.cone_cos_inner} for a in al])
        env = g("environment").view(np.float32)
        s.set_constant_environment(env[:3], float(env[3]))
        recs = g("textures").view(np.int32).reshape(-1, 6) if "textures" in sections else np.zeros((0, 6), np.int32)
        texels = g("texels").view(np.float32) if "texels" in sections else np.zeros(0, np.float32)
        off = 0
        for w, h, filt, ws, wt, _ in recs:
            n = int(w) * int(h) * 4
            s.add_texture(texels[off:off + n].reshape(int(h), int(w), 4), int(filt), int(ws), int(wt))
            off += n
        if "env_map" in sections:
            em = g("env_map").view(np.int32)
            if em[0]:
                s.set_environment_map(int(em[1]), env[:3], float(env[3]), float(em[2:3].view(np.float32)[0]))
        s.build_accel()
        st = g("settings").view(np.int64)
        return s, {"width": int(st[0]), "height": int(st[1]), "spp": int(st[2]), "max_depth": int(st[3]),
                   "sampler": int(st[4]), "seed": int(st[5])}

    def close(self):
        if self._h:
            self._lib.yrhost_scene_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- queries ----------------------------------------------------------
    @property
    def info(self) -> N.SceneInfo:
        info = N.SceneInfo()
        self._lib.yrhost_scene_info_get(self._h, C.byref(info))
        return info

    def geometry(self):
        """(vertices, indices, primitives) numpy copies."""
        info = self.info
        v, i, p = C.c_void_p(), C.c_void_p(), C.c_void_p()
        self._lib.yrhost_scene_geometry(self._h, C.byref(v), C.byref(i), C.byref(p))
        return (_view(v.value, VERTEX_DTYPE, info.vertex_count), _view(i.value, "<u4", info.index_count),
                _view(p.value, PRIMITIVE_DTYPE, info.primitive_count))

    def build_accel(self, threads: int = 0) -> float:
        secs = C.c_double()
        rc = self._lib.yrhost_build_accel(self._h, threads, C.byref(secs))
        if rc != 0:
            raise YrhError(self._lib.yrhost_scene_error(self._h).decode())
        return secs.value

    def flat_bvh(self) -> BvhArrays:
        d = N.BvhDesc()
        if self._lib.yrhost_accel_flat(self._h, C.byref(d)) != 0:
            raise YrhError("no acceleration structure")
        return BvhArrays.from_desc(d)

    def blas(self, index: int) -> BvhArrays:
        d = N.BvhDesc()
        if self._lib.yrhost_accel_blas(self._h, index, C.byref(d)) != 0:
            raise YrhError(f"no BLAS {index}")
        return BvhArrays.from_desc(d)

    def tlas(self):
        """(tlas BvhArrays (no primitives), tlas primitives, instances)."""
        d = N.BvhDesc()
        prims, nprims, inst, ninst = C.c_void_p(), C.c_int32(), C.c_void_p(), C.c_int32()
        if self._lib.yrhost_accel_tlas(self._h, C.byref(d), C.byref(prims), C.byref(nprims), C.byref(inst),
                                       C.byref(ninst)) != 0:
            raise YrhError("no acceleration structure")
        t = BvhArrays(_view(d.nodes, NODE_DTYPE, d.node_count), _view(d.primitive_indices, "<i4",
                                                                      d.primitive_index_count),
                      np.zeros(0, PRIM_DTYPE), d.max_depth)
        return t, _view(prims.value, TLAS_PRIM_DTYPE, nprims.value), _view(inst.value, INSTANCE_DTYPE, ninst.value)

    def camera_rays(self, width: int, height: int, t_min=1e-5, t_max=float("inf")):
        org = np.empty((width * height, 4), np.float32)
        dirs = np.empty((width * height, 4), np.float32)
        rc = self._lib.yrhost_camera_rays(self._h, width, height, t_min, t_max, _ptr(org), _ptr(dirs))
        if rc != 0:
            raise YrhError("yrhost_camera_rays failed")
        return org, dirs

    # -- edits for in-code test jobs ---------------------------------------------
    def set_camera(self, origin, forward, right, up, fov_y_radians):
        c = np.array([*origin, *forward, *right, *up, fov_y_radians], np.float32)
        self._lib.yrhost_scene_set_camera(self._h, c.ctypes.data)

    def set_materials(self, materials):
        """materials: list of dicts with keys of native.Material (kind, reflectance, emission, ...)."""
        arr = material_array(materials)
        if self._lib.yrhost_scene_set_materials(self._h, C.cast(arr, C.c_void_p), len(materials)) != 0:
            raise YrhError("set_materials failed")

    def add_texture(self, texels, filter=N.FILTER_BILINEAR, wrap_s=N.WRAP_REPEAT, wrap_t=None) -> int:
        """Appends an RGBA texture (texels: height x width x 4 linear floats, row 0 first;
        filter / wraps: native.FILTER_* / WRAP_*; wrap_t defaults to wrap_s) to the
        scene's texture list and returns its index, which materials name in
        reflectance_texture / alpha_texture / normal_map."""
        t = np.ascontiguousarray(texels, np.float32)
        if t.ndim != 3 or t.shape[2] != 4:
            raise ValueError("texels must be height x width x 4")
        d = N.Texture()
        d.height, d.width = int(t.shape[0]), int(t.shape[1])
        d.filter = int(filter)
        d.wrap_s = int(wrap_s)
        d.wrap_t = int(wrap_s if wrap_t is None else wrap_t)
        d.texels = _ptr(t)
        rc = self._lib.yrhost_scene_add_texture(self._h, C.byref(d))
        if rc < 0:
            raise YrhError("add_texture failed: " + self._lib.yrhost_scene_error(self._h).decode())
        return rc

    def set_environment_map(self, texture_index, radiance=(1.0, 1.0, 1.0), strength=1.0, rotation_radians=0.0):
        """An active equirect environment on texture `texture_index` (bilinear, repeat)."""
        r = (C.c_float * 3)(*radiance)
        if self._lib.yrhost_scene_set_environment_map(self._h, int(texture_index), r, float(strength),
                                                      float(rotation_radians)) != 0:
            raise YrhError("set_environment_map failed: " + self._lib.yrhost_scene_error(self._h).decode())

    def set_spheres(self, spheres):
        """spheres: list of (center (3), radius, material_index[, flip_normals]) (yr::RenderSphere);
        the acceleration structure must be built again."""
        arr = (N.Sphere * max(1, len(spheres)))()
        for k, sp in enumerate(spheres):
            arr[k].center[:] = [float(c) for c in sp[0]]
            arr[k].radius = float(sp[1])
            arr[k].material_index = int(sp[2])
            arr[k].area_light_index = -1
            arr[k].flip_normals = 1 if (len(sp) > 3 and sp[3]) else 0
        if self._lib.yrhost_scene_set_spheres(self._h, C.cast(arr, C.c_void_p), len(spheres)) != 0:
            raise YrhError("set_spheres failed")

    LIGHT_KINDS = {"point": 0, "spot": 1, "distant": 2}

    def set_analytic_lights(self, lights):
        """lights: list of dicts (yr::AnalyticLight): kind ("point" | "spot" |
        "distant"), position, direction, intensity (rgb), and for spot lights
        cone_angle / cone_cos_inner (cosines of the outer / inner cone). The
        selection tables (power alias table, light tree) are rebuilt."""
        arr = (N.AnalyticLight * max(1, len(lights)))()
        for k, l in enumerate(lights):
            arr[k].kind = self.LIGHT_KINDS[l["kind"]] if isinstance(l["kind"], str) else int(l["kind"])
            arr[k].position[:] = [float(v) for v in l.get("position", (0.0, 0.0, 0.0))]
            arr[k].direction[:] = [float(v) for v in l.get("direction", (0.0, 0.0, 0.0))]
            arr[k].intensity[:] = [float(v) for v in l["intensity"]]
            arr[k].cone_angle = float(l.get("cone_angle", 0.0))
            arr[k].cone_cos_inner = float(l.get("cone_cos_inner", 0.0))
        if self._lib.yrhost_scene_set_analytic_lights(self._h, C.cast(arr, C.c_void_p), len(lights)) != 0:
            raise YrhError("set_analytic_lights failed")

    def spheres(self) -> np.ndarray:
        """The scene's spheres as a SPHERE_DTYPE array (a copy)."""
        p, n = C.c_void_p(), C.c_int32()
        self._lib.yrhost_scene_spheres(self._h, C.byref(p), C.byref(n))
        if n.value == 0:
            return np.zeros(0, SPHERE_DTYPE)
        return _view(p.value, SPHERE_DTYPE, n.value)

    def set_emissives(self, emissives):
        """emissives: list of (primitive_index, (r, g, b), area)."""
        arr = (N.Emissive * max(1, len(emissives)))()
        for k, (prim, rad, area) in enumerate(emissives):
            arr[k].primitive_index = prim
            arr[k].radiance[:] = rad
            arr[k].area = area
        self._lib.yrhost_scene_set_emissives(self._h, C.cast(arr, C.c_void_p), len(emissives))

    def set_constant_environment(self, radiance, strength=1.0):
        r = np.array(radiance, np.float32)
        self._lib.yrhost_scene_set_constant_environment(self._h, r.ctypes.data, strength)

    def geometry_desc(self) -> N.GeometryDesc:
        """The C-ABI geometry descriptor (pointers into this scene; keep the scene alive)."""
        g = N.GeometryDesc()
        if self._lib.yrhost_geometry_desc(self._h, C.byref(g)) != 0:
            raise YrhError("no acceleration structure")
        return g

    def shading_desc(self) -> N.ShadingDesc:
        d = N.ShadingDesc()
        if self._lib.yrhost_shading_desc(self._h, C.byref(d)) != 0:
            raise YrhError(self._lib.yrhost_scene_error(self._h).decode())
        return d

    def render_desc(self, width, height, spp, max_depth=8, sampler=N.SAMPLER_OWEN_SOBOL, seed=0x59414F524159,
                    radiance_clamp=0.0, sample_index=0) -> N.RenderDesc:
        d = N.RenderDesc()
        if self._lib.yrhost_render_desc(self._h, width, height, sampler, spp, max_depth, seed, radiance_clamp,
                                        sample_index, C.byref(d)) != 0:
            raise YrhError("yrhost_render_desc failed")
        return d

    def backend_render(self, width, height, spp, max_depth=8, sampler=N.SAMPLER_OWEN_SOBOL, seed=0x59414F524159,
                       radiance_clamp=0.0, device=0, requested_backend=0, resume_film=None,
                       resume_completed_spp=0, stop_before_start=False, cancel_after_passes=0, cancel_error=None,
                       adaptive=None):
        """yr::HipBackend Prepare + Render (the reference's RenderBackend flow); adaptive =
        (min_spp, relative_error, absolute_error, confidence) turns on adaptive sampling.

        Returns (report dict, film or None); the film is a FILM32_DTYPE
        (yr::FilmPixel) array of shape (height, width)."""
        st = N.RenderSettings(width, height, sampler, spp, max_depth, radiance_clamp, seed, device,
                              requested_backend)
        if adaptive is not None:
            st.adaptive_enabled = 1
            st.adaptive_min_spp, st.adaptive_relative_error, st.adaptive_absolute_error, st.adaptive_confidence = adaptive
        rq = N.RenderRequest()
        keep = None
        if resume_film is not None:
            keep = np.ascontiguousarray(resume_film, dtype=FILM32_DTYPE)
            rq.resume_film = keep.ctypes.data
            rq.resume_height, rq.resume_width = keep.shape
        rq.resume_completed_spp = resume_completed_spp
        rq.stop_before_start = 1 if stop_before_start else 0
        rq.cancel_after_passes = cancel_after_passes
        rq.cancel_error = cancel_error.encode() if cancel_error else None
        cap = max(0, width) * max(0, height)
        film = np.zeros(cap, FILM32_DTYPE)
        rep = N.RenderReport()
        rc = self._lib.yrhost_backend_render(self._h, C.byref(st), C.byref(rq), film.ctypes.data if cap else None,
                                             cap, C.byref(rep))
        if rc != 0:
            raise YrhError(f"yrhost_backend_render rejected its arguments ({rc})")
        report = {k: getattr(rep, k) for k, _ in N.RenderReport._fields_ if k != "error"}
        report["error"] = rep.error.decode()
        del keep
        if not rep.film_valid:
            return report, None
        return report, film[: rep.film_width * rep.film_height].reshape(rep.film_height, rep.film_width)

    @property
    def handle(self):
        return self._h


def uneven_cameras(count=32, seed=5, min_pixels=128 * 128, max_pixels=1920 * 1080, min_log2_spp=2,
                   max_log2_spp=8, distance=(25.0, 50.0), target_spread=5.0):
    """Cameras of uneven size and spp for the batched pipeline: pixel counts
    log-uniform in [min_pixels, max_pixels] at 16:9 (clamped), spp a power of two
    2^k with k uniform in [min_log2_spp, max_log2_spp], eyes on a ring `distance`
    out and 3..15 up, looking at points within target_spread of the centre, fov
    30..60 degrees. Returns [(width, height, spp, (eye, forward, right, up, fov))]."""
    rng = np.random.default_rng(seed)
    max_w = int(round(np.sqrt(max_pixels * 16.0 / 9.0)))
    min_w = int(round(np.sqrt(min_pixels)))
    out = []
    for _ in range(count):
        n = float(np.exp(rng.uniform(np.log(min_pixels), np.log(max_pixels))))
        w = int(np.clip(round(np.sqrt(n * 16.0 / 9.0)), min_w, max_w))
        h = int(np.clip(round(w * 9.0 / 16.0), max(1, (min_w * 9) // 16), (max_w * 9) // 16 + 1))
        spp = int(2 ** rng.integers(min_log2_spp, max_log2_spp + 1))
        az, dist, el = rng.uniform(0, 2 * np.pi), rng.uniform(*distance), rng.uniform(3, 15)
        eye = np.array([dist * np.sin(az), el, dist * np.cos(az)])
        target = np.array([rng.uniform(-target_spread, target_spread), 1.0,
                           rng.uniform(-target_spread, target_spread)])
        fov = rng.uniform(30, 60)
        f = (target - eye) / np.linalg.norm(target - eye)
        r = np.cross(f, [0.0, 1.0, 0.0])
        r /= np.linalg.norm(r)
        u = np.cross(r, f)
        out.append((w, h, spp, (eye, f, r, u, np.radians(fov))))
    return out


def random_rays(count: int, seed: int = 3, t_min=1e-5, t_max=float("inf")):
    org = np.empty((count, 4), np.float32)
    dirs = np.empty((count, 4), np.float32)
    rc = N.host().yrhost_random_rays(count, seed, t_min, t_max, _ptr(org) if count else None,
                                     _ptr(dirs) if count else None)
    if rc != 0:
        raise YrhError("yrhost_random_rays failed")
    return org, dirs


class DeviceBuffer:
    def __init__(self, device: "Device", nbytes: int):
        self.device = device
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        device._check(device._lib.yrh_device_alloc(device._ctx, max(self.nbytes, 16), C.byref(p)), "alloc")
        self.ptr = p.value

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        self.device._check(self.device._lib.yrh_copy_to_device(self.device._ctx, self.ptr, _ptr(a), a.nbytes),
                           "copy_to_device")
        return self

    def download(self, dtype, shape) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes <= self.nbytes
        self.device._check(self.device._lib.yrh_copy_to_host(self.device._ctx, _ptr(out), self.ptr, out.nbytes),
                           "copy_to_host")
        return out

    def free(self):
        if self.ptr:
            self.device._lib.yrh_device_free(self.device._ctx, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            if self.ptr and self.device._ctx:
                self.free()
        except Exception:
            pass


class Device:
    """One yrh_context: a GPU, a HIP stream and the uploaded scene."""

    def __init__(self, device: int = 0):
        self._lib = N.yrh()
        ctx = C.c_void_p()
        rc = self._lib.yrh_create(device, C.byref(ctx))
        if rc != 0:
            raise YrhError(f"yrh_create({device}) failed with {rc} (no HIP device?)")
        self._ctx = ctx

    def _check(self, rc, what):
        if rc != 0:
            msg = self._lib.yrh_last_error(self._ctx)
            raise YrhError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    def close(self):
        if self._ctx:
            self._lib.yrh_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def ctx(self):
        return self._ctx

    def info(self):
        name = C.create_string_buffer(256)
        cus = C.c_int()
        self._check(self._lib.yrh_device_info(self._ctx, name, 256, C.byref(cus)), "device_info")
        return name.value.decode(), cus.value

    def synchronize(self):
        self._check(self._lib.yrh_synchronize(self._ctx), "synchronize")

    def alloc(self, nbytes) -> DeviceBuffer:
        return DeviceBuffer(self, nbytes)

    def to_device(self, a: np.ndarray) -> DeviceBuffer:
        a = np.ascontiguousarray(a)
        return DeviceBuffer(self, a.nbytes).upload(a)

    def upload_scene(self, scene: Scene):
        self._check(N.host().yrhost_upload_geometry(scene.handle, self._ctx), "upload_geometry")

    def upload_geometry(self, desc: N.GeometryDesc):
        self._check(self._lib.yrh_upload_geometry(self._ctx, C.byref(desc)), "upload_geometry")

    def set_trace_lds_levels(self, levels: int):
        self._check(self._lib.yrh_set_trace_lds_levels(self._ctx, levels), "set_trace_lds_levels")

    def set_trace_timing(self, enable: bool):
        self._check(self._lib.yrh_set_trace_timing(self._ctx, int(bool(enable))), "set_trace_timing")

    def set_render_overlap(self, enable: bool):
        """Shadow rays of a bounce on a second stream beside the next bounce's traversal."""
        self._check(self._lib.yrh_set_render_overlap(self._ctx, int(bool(enable))), "set_render_overlap")

    def trace_phase_ms(self):
        """(ordering ms, traversal ms) of the last timed ray-cast call (waits for it)."""
        a, b = C.c_float(), C.c_float()
        self._check(self._lib.yrh_trace_phase_ms(self._ctx, C.byref(a), C.byref(b)), "trace_phase_ms")
        return a.value, b.value

    def set_trace_order(self, order: int):
        """N.TRACE_ORDER_ORIGIN (default) or N.TRACE_ORDER_INPUT for the ray-cast calls."""
        self._check(self._lib.yrh_set_trace_order(self._ctx, order), "set_trace_order")

    def set_trace_config(self, blocks_per_cu=0, block_size=0):
        self._check(self._lib.yrh_set_trace_config(self._ctx, blocks_per_cu, block_size), "set_trace_config")

    # -- timing on the context's stream ---------------------------------------
    def event(self):
        e = C.c_void_p()
        self._check(self._lib.yrh_event_create(self._ctx, C.byref(e)), "event_create")
        return e.value

    def record(self, event):
        self._check(self._lib.yrh_event_record(self._ctx, event), "event_record")

    def elapsed_ms(self, start, stop) -> float:
        ms = C.c_float()
        self._check(self._lib.yrh_event_elapsed_ms(self._ctx, start, stop, C.byref(ms)), "event_elapsed")
        return ms.value

    def destroy_event(self, event):
        self._lib.yrh_event_destroy(self._ctx, event)

    # -- device-resident launches ------------------------------------------
    def trace_closest_device(self, org: DeviceBuffer, dirs: DeviceBuffer, n: int, hit: DeviceBuffer,
                             instance: DeviceBuffer | None = None, node_tests: DeviceBuffer | None = None,
                             tri_tests: DeviceBuffer | None = None, sphere_tests: DeviceBuffer | None = None):
        rays = N.RaySoa(org.ptr, dirs.ptr)
        hits = N.HitSoa(hit.ptr, instance.ptr if instance else None, node_tests.ptr if node_tests else None,
                        tri_tests.ptr if tri_tests else None, sphere_tests.ptr if sphere_tests else None)
        self._check(self._lib.yrh_trace_closest(self._ctx, C.byref(rays), n, C.byref(hits)), "trace_closest")

    def trace_any_device(self, org: DeviceBuffer, dirs: DeviceBuffer, n: int, occluded: DeviceBuffer,
                         node_tests: DeviceBuffer | None = None, tri_tests: DeviceBuffer | None = None):
        rays = N.RaySoa(org.ptr, dirs.ptr)
        self._check(self._lib.yrh_trace_any(self._ctx, C.byref(rays), n, occluded.ptr,
                                            node_tests.ptr if node_tests else None,
                                            tri_tests.ptr if tri_tests else None), "trace_any")

    # -- wavefront path tracing ----------------------------------------------
    def upload_shading(self, scene: Scene):
        self._check(N.host().yrhost_upload_shading(scene.handle, self._ctx), "upload_shading")

    def film_reset(self, width: int, height: int):
        self._check(self._lib.yrh_film_reset(self._ctx, width, height), "film_reset")

    def render_pass(self, desc: N.RenderDesc):
        self._check(self._lib.yrh_render_pass(self._ctx, C.byref(desc)), "render_pass")

    def film_read(self, width: int, height: int) -> np.ndarray:
        out = np.zeros(width * height, FILM64_DTYPE)
        self._check(self._lib.yrh_film_read(self._ctx, out.ctypes.data, len(out)), "film_read")
        return out

    def film_load(self, pixels: np.ndarray):
        pixels = np.ascontiguousarray(pixels, FILM64_DTYPE)
        self._check(self._lib.yrh_film_load(self._ctx, pixels.ctypes.data, len(pixels)), "film_load")

    def render_stats(self) -> dict:
        s = N.RenderStats()
        self._check(self._lib.yrh_render_stats_read(self._ctx, C.byref(s)), "render_stats_read")
        return {k: int(getattr(s, k)) for k, _ in N.RenderStats._fields_}

    def render(self, scene: Scene, width, height, spp, max_depth=8, sampler=N.SAMPLER_OWEN_SOBOL,
               seed=0x59414F524159, radiance_clamp=0.0, first_sample=0, upload=True, batch=1):
        """Uploads the scene (unless upload=False), renders samples [first_sample, spp), `batch`
        sample indices per pass, and returns (FP64 film records, stats)."""
        if upload:
            self.upload_scene(scene)
            self.upload_shading(scene)
        self.film_reset(width, height)
        for s in range(first_sample, spp, batch):
            d = scene.render_desc(width, height, spp, max_depth, sampler, seed, radiance_clamp, s)
            d.sample_count = min(batch, spp - s)
            self.render_pass(d)
        self.synchronize()
        return self.film_read(width, height), self.render_stats()

    def film_update_active(self, spp: int, adaptive=None):
        """Writes the device list of active pixels (fewer than spp samples and, with
        adaptive = (min_spp, relative_error, absolute_error, confidence), not converged)
        for the next pixel-list pass; returns (active pixels, converged pixels)."""
        a = N.AdaptiveDesc(0, 0, 0.0, 0.0, 0.0) if adaptive is None else N.AdaptiveDesc(1, *adaptive)
        active, converged = C.c_int64(), C.c_int64()
        self._check(self._lib.yrh_film_update_active(self._ctx, C.byref(a), spp, C.byref(active), C.byref(converged)),
                    "film_update_active")
        return active.value, converged.value

    def render_adaptive(self, scene: Scene, width, height, spp, adaptive, max_depth=8, sampler=N.SAMPLER_OWEN_SOBOL,
                        seed=0x59414F524159, radiance_clamp=0.0, upload=True, batch=1):
        """RenderCpuPathTrace's active-pixel loop on the device: each pass renders the next
        sample of every active pixel, until none is active or spp passes ran. With batch > 1
        one device pass stands for `batch` of them (each pixel's samples are added in order
        until it converges, so the film is the same; the counters include the samples not
        added); the first covers max(batch, min_spp) passes, before which no pixel can
        converge. Returns (FP64 film records, stats, passes, converged pixels)."""
        if upload:
            self.upload_scene(scene)
            self.upload_shading(scene)
        self.film_reset(width, height)
        active, converged = self.film_update_active(spp, adaptive)
        passes = 0
        while passes < spp and active > 0:
            first = max(batch, adaptive[0]) if (passes == 0 and adaptive is not None and batch > 1) else batch
            d = scene.render_desc(width, height, spp, max_depth, sampler, seed, radiance_clamp, passes)
            d.pixel_list = 1
            d.sample_count = min(first, spp - passes)
            self.render_pass(d)
            active, converged = self.film_update_active(spp, adaptive)
            passes += d.sample_count
        return self.film_read(width, height), self.render_stats(), passes, converged

    def suggest_paths_per_pass(self, max_depth: int = 8) -> int:
        """Paths one render pass should hold for the free device memory (yrh_suggest_paths_per_pass)."""
        v = C.c_int64()
        self._check(self._lib.yrh_suggest_paths_per_pass(self._ctx, max_depth, C.byref(v)), "suggest_paths_per_pass")
        return v.value

    def sample_values(self, kind, seed, spp, x, y, sample, dim, dims):
        n = len(x)
        arrs = [np.ascontiguousarray(a, np.int32) for a in (x, y, sample, dim)]
        out = np.zeros(n * dims, np.float32)
        self._check(self._lib.yrh_sample_values(self._ctx, kind, seed, spp, n, *[a.ctypes.data for a in arrs], dims,
                                                out.ctypes.data), "sample_values")
        return out.reshape(n, dims) if dims == 2 else out

    def bsdf_query(self, materials, material_index, queries):
        """Device BSDF queries: materials = list of material dicts (Scene.set_materials keys),
        queries = (n, 12) float32 rows (wo, wi, normal, u0, u1, unused). Returns (n, 12) rows
        (f, pdf, sampled wi, weight, pdf, flags = valid + 2 specular + 4 delta)."""
        arr = material_array(materials)
        idx = np.ascontiguousarray(material_index, np.int32)
        q = np.ascontiguousarray(queries, np.float32).reshape(-1, 12)
        out = np.zeros((len(q), 12), np.float32)
        self._check(self._lib.yrh_bsdf_query(self._ctx, C.cast(arr, C.c_void_p), len(materials), len(q),
                                             idx.ctypes.data, q.ctypes.data, out.ctypes.data), "bsdf_query")
        return out

    def texture_query(self, texels, uv, footprint=None, filter=N.FILTER_BILINEAR, wrap_s=N.WRAP_REPEAT, wrap_t=None):
        """Device SampleTexture4 of an RGBA texture (height x width x 4) at uv (n x 2) with
        footprints (n x 4: dudx, dvdx, dudy, dvdy) or none; returns n x 4."""
        t = np.ascontiguousarray(texels, np.float32)
        d = N.Texture()
        d.height, d.width = (int(t.shape[0]), int(t.shape[1])) if t.size else (0, 0)
        d.filter, d.wrap_s = int(filter), int(wrap_s)
        d.wrap_t = int(wrap_s if wrap_t is None else wrap_t)
        d.texels = t.ctypes.data if t.size else None
        q = np.ascontiguousarray(uv, np.float32).reshape(-1, 2)
        fp = None if footprint is None else np.ascontiguousarray(footprint, np.float32).reshape(-1, 4)
        out = np.zeros((len(q), 4), np.float32)
        self._check(self._lib.yrh_texture_query(self._ctx, C.byref(d), len(q), q.ctypes.data,
                                                None if fp is None else fp.ctypes.data, out.ctypes.data),
                    "texture_query")
        return out

    def camera_rays(self, desc: N.RenderDesc):
        n = desc.width * desc.height
        d_org, d_dir = self.alloc(16 * n), self.alloc(16 * n)
        self._check(self._lib.yrh_generate_camera_rays(self._ctx, C.byref(desc), d_org.ptr, d_dir.ptr), "camera_rays")
        self.synchronize()
        return d_org.download(np.float32, (n, 4)), d_dir.download(np.float32, (n, 4))

    # -- batched multi-camera rendering --------------------------------------
    def render_batch(self, jobs, max_depth=8, sampler=N.SAMPLER_OWEN_SOBOL, seed=0x59414F524159, radiance_clamp=0.0,
                     paths_per_pass=0, use_graph=True):
        """jobs: list of RenderDesc (camera, tan_half_fov, width, height, spp; e.g. from Scene.render_desc).
        Returns (FP64 film records per job, job completion ms, passes, stats)."""
        arr = (N.BatchJob * len(jobs))()
        for k, j in enumerate(jobs):
            b = arr[k]
            b.camera_origin[:] = list(j.camera_origin)
            b.camera_forward[:] = list(j.camera_forward)
            b.camera_right[:] = list(j.camera_right)
            b.camera_up[:] = list(j.camera_up)
            b.tan_half_fov = j.tan_half_fov
            b.width, b.height, b.spp = j.width, j.height, j.spp
        d = N.BatchDesc(C.cast(arr, C.c_void_p), len(jobs), sampler, max_depth, radiance_clamp, seed, paths_per_pass,
                        1 if use_graph else 0, 0)
        ms = np.zeros(len(jobs), np.float32)
        passes = C.c_int32()
        self._check(self._lib.yrh_batch_render(self._ctx, C.byref(d), ms.ctypes.data, C.byref(passes)), "render_batch")
        films = []
        for k, j in enumerate(jobs):
            n = j.width * j.height
            f = np.zeros(n, FILM64_DTYPE)
            self._check(self._lib.yrh_batch_film_read(self._ctx, k, f.ctypes.data, n), "batch_film_read")
            films.append(f)
        return films, ms, passes.value, self.render_stats()

    # -- learned guiding -----------------------------------------------------
    def guide_configure(self, bounds_min, bounds_max, log2_table_size=16, min_resolution=8, max_resolution=256,
                        learning_rate=1e-2, guide_fraction=0.5, sample_capacity=1 << 20, seed=7):
        d = N.GuideDesc((C.c_float * 3)(*bounds_min), (C.c_float * 3)(*bounds_max), log2_table_size, min_resolution,
                        max_resolution, learning_rate, guide_fraction, sample_capacity, seed)
        self._check(self._lib.yrh_guide_configure(self._ctx, C.byref(d)), "guide_configure")
        return d

    def guide_set_mode(self, mode: int):
        self._check(self._lib.yrh_guide_set_mode(self._ctx, mode), "guide_set_mode")

    def guide_set_cell_budget(self, nbytes: int):
        """Device memory sampling passes may spend on per-cell grid copies (0 = off)."""
        self._check(self._lib.yrh_guide_set_cell_budget(self._ctx, nbytes), "guide_set_cell_budget")

    def guide_set_depth_limit(self, depths: int):
        """Guide only the first `depths` bounces (0 = every bounce that may continue)."""
        self._check(self._lib.yrh_guide_set_depth_limit(self._ctx, depths), "guide_set_depth_limit")

    def guide_train(self, steps: int, batch: int):
        self._check(self._lib.yrh_guide_train(self._ctx, steps, batch), "guide_train")

    def guide_add_samples(self, samples: np.ndarray):
        smp = np.ascontiguousarray(samples, np.float32).reshape(-1, 13)
        self._check(self._lib.yrh_guide_add_samples(self._ctx, len(smp), smp.ctypes.data), "guide_add_samples")

    def guide_read_samples(self, n: int) -> np.ndarray:
        out = np.zeros((n, 13), np.float32)
        got = C.c_int64()
        self._check(self._lib.yrh_guide_read_samples(self._ctx, n, out.ctypes.data, C.byref(got)), "guide_read_samples")
        return out[:got.value]

    def guide_sample_stats(self):
        n, w = C.c_int64(), C.c_float()
        self._check(self._lib.yrh_guide_sample_stats(self._ctx, C.byref(n), C.byref(w)), "guide_sample_stats")
        return n.value, w.value

    def guide_info(self):
        m, g = C.c_int32(), C.c_int64()
        res, dense = np.zeros(8, np.int32), np.zeros(8, np.int32)
        self._check(self._lib.yrh_guide_info(self._ctx, C.byref(m), C.byref(g), res.ctypes.data, dense.ctypes.data),
                    "guide_info")
        return m.value, g.value, res, dense

    def guide_params(self):
        m, g, _, _ = self.guide_info()
        mlp, grid = np.zeros(m, np.float32), np.zeros(g, np.float32)
        self._check(self._lib.yrh_guide_params_read(self._ctx, mlp.ctypes.data, grid.ctypes.data), "guide_params")
        return mlp, grid

    def guide_set_params(self, mlp=None, grid=None):
        m = None if mlp is None else np.ascontiguousarray(mlp, np.float32)
        g = None if grid is None else np.ascontiguousarray(grid, np.float32)
        self._check(self._lib.yrh_guide_params_write(self._ctx, None if m is None else m.ctypes.data,
                                                      None if g is None else g.ctypes.data), "guide_set_params")

    def guide_query(self, rows: np.ndarray) -> np.ndarray:
        """rows: (n, 9) position, normal, outgoing -> (n, 40) mixtures."""
        r = np.ascontiguousarray(rows, np.float32).reshape(-1, 9)
        out = np.zeros((len(r), 40), np.float32)
        self._check(self._lib.yrh_guide_query(self._ctx, len(r), r.ctypes.data, out.ctypes.data), "guide_query")
        return out

    def guide_gradients(self, samples: np.ndarray, w_norm: float):
        """samples: (n, 13) position, normal, outgoing, direction, weight -> (grad_mlp, grad_grid, loss)."""
        smp = np.ascontiguousarray(samples, np.float32).reshape(-1, 13)
        m, g, _, _ = self.guide_info()
        gm, gg, loss = np.zeros(m, np.float32), np.zeros(g, np.float32), C.c_float()
        self._check(self._lib.yrh_guide_gradients(self._ctx, len(smp), smp.ctypes.data, w_norm, gm.ctypes.data,
                                                   gg.ctypes.data, C.byref(loss)), "guide_gradients")
        return gm, gg, loss.value

    def guide_mixture_eval(self, mixture: np.ndarray, dirs: np.ndarray, u: np.ndarray):
        m = np.ascontiguousarray(mixture, np.float32).reshape(-1, 40)
        d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
        uu = np.ascontiguousarray(u, np.float32).reshape(-1, 3)
        pdf, smp = np.zeros(len(m), np.float32), np.zeros((len(m), 4), np.float32)
        self._check(self._lib.yrh_guide_mixture_eval(self._ctx, len(m), m.ctypes.data, d.ctypes.data, uu.ctypes.data,
                                                      pdf.ctypes.data, smp.ctypes.data), "guide_mixture_eval")
        return pdf, smp

    # -- host convenience ---------------------------------------------------
    def trace_closest(self, org: np.ndarray, dirs: np.ndarray, counters: bool = False):
        """Returns dict(t, u, v, triangle, instance[, node_tests, triangle_tests, sphere_tests]) as
        numpy arrays; `triangle` is the hit id (-1 miss, >= 0 flat triangle index, -2 - s sphere s)."""
        n = len(org)
        d_org, d_dir = self.to_device(org.astype(np.float32)), self.to_device(dirs.astype(np.float32))
        d_hit = self.alloc(16 * n)
        d_inst = self.alloc(4 * n)
        d_nt = self.alloc(4 * n) if counters else None
        d_tt = self.alloc(4 * n) if counters else None
        d_st = self.alloc(4 * n) if counters else None
        self.trace_closest_device(d_org, d_dir, n, d_hit, d_inst, d_nt, d_tt, d_st)
        self.synchronize()
        hit = d_hit.download(np.float32, (n, 4))
        out = {"t": hit[:, 0].copy(), "u": hit[:, 1].copy(), "v": hit[:, 2].copy(),
               "triangle": hit[:, 3].copy().view(np.int32), "instance": d_inst.download(np.int32, (n,))}
        if counters:
            out["node_tests"] = d_nt.download(np.uint32, (n,))
            out["triangle_tests"] = d_tt.download(np.uint32, (n,))
            out["sphere_tests"] = d_st.download(np.uint32, (n,))
        return out

    def trace_any(self, org: np.ndarray, dirs: np.ndarray, counters: bool = False):
        n = len(org)
        d_org, d_dir = self.to_device(org.astype(np.float32)), self.to_device(dirs.astype(np.float32))
        d_occ = self.alloc(4 * n)
        d_nt = self.alloc(4 * n) if counters else None
        d_tt = self.alloc(4 * n) if counters else None
        self.trace_any_device(d_org, d_dir, n, d_occ, d_nt, d_tt)
        self.synchronize()
        out = {"occluded": d_occ.download(np.uint32, (n,))}
        if counters:
            out["node_tests"] = d_nt.download(np.uint32, (n,))
            out["triangle_tests"] = d_tt.download(np.uint32, (n,))
        return out
