"""Writes the golden fixtures of tests/golden/ from the reference itself.

Run in a container where the reference's sources are present and
oracle/_ref/libyrref.so is built from them (`make -C oracle ref`):

    python tests/golden/make_golden.py [fixture names]   # default: all

Each fixture holds its inputs (or, for the procedural Cornell workload, a
digest of them) and the reference's outputs on them, so tests can pin the
oracle and the device path where the reference is absent. Sizes follow the
reference-only pins of the traversal and path-tracing path: SAH node arrays
(src/accel/bvh_split_policy.cpp:105-112 order), per-ray BvhHit + node /
triangle counters (src/accel/bvh_traversal.cpp:253-345,
src/accel/two_level_bvh.cpp:365-481), yr::Sampler bits
(src/sampling/sampler.cpp) and FP32 films + stats of RenderCpuPathTrace
(src/backends/cpu/cpu_path_tracer.cpp), plain and with adaptive sampling
(its active-pixel loop and Film::IsConverged, src/film/film.cpp:113-121),
BSDF eval / pdf / sample tuples (src/shading/bsdf.cpp:17-329), and hits,
counters and a film of scenes with analytic spheres
(src/geometry/intersection.cpp:108-151, src/integrators/surface_query.cpp:129-147).
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle_py as O  # noqa: E402
from tests.helpers import MeshArrays  # noqa: E402
from yaoray_amd import Scene, random_rays  # noqa: E402

SEED = 0x59414F524159


def soup(count, seed, edge_min, edge_max):
    """Random triangles: centres uniform in the unit cube, edge lengths
    log-uniform in [edge_min, edge_max], random orientation (the ray-cast
    workload's rule, with longer edges so a few thousand triangles still
    stop a good share of the rays)."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(0.0, 1.0, (count, 1, 3))
    edge = np.exp(rng.uniform(np.log(edge_min), np.log(edge_max), (count, 1, 1)))
    offs = rng.normal(size=(count, 3, 3))
    offs /= np.linalg.norm(offs, axis=2, keepdims=True)
    return (c + 0.5 * edge * offs).astype(np.float32)


def geometry_digest(scene):
    v, i, p = scene.geometry()
    h = hashlib.sha256()
    for a in (v, i, p):
        h.update(np.ascontiguousarray(a).tobytes())
    h.update(np.array(list(scene.info.camera), np.float32).tobytes())
    return h.hexdigest()


def sampler_fixture():
    rng = np.random.default_rng(101)
    n = 1024
    x = rng.integers(0, 2048, n).astype(np.int32)
    y = rng.integers(0, 2048, n).astype(np.int32)
    sample = rng.integers(0, 16, n).astype(np.int32)
    dim = rng.integers(0, 48, n).astype(np.int32)
    out = {"x": x, "y": y, "sample": sample, "dim": dim, "seed": np.uint64(SEED)}
    for kind in range(4):
        for spp in (1, 13, 16):
            s = np.minimum(sample, spp - 1).astype(np.int32)
            out[f"v1_{kind}_{spp}"] = O.ref_sample_values(kind, SEED, spp, x, y, s, dim, 1)
            out[f"v2_{kind}_{spp}"] = O.ref_sample_values(kind, SEED, spp, x, y, s, dim, 2)
    return out


def trace_outputs(r):
    hit = np.stack([r["t"], r["u"], r["v"], r["triangle"].view(np.float32)], axis=1)
    return {"hit": hit, "instance": r["instance"], "node_tests": r["node_tests"],
            "triangle_tests": r["triangle_tests"]}


def flat_fixture():
    tris = soup(3000, 202, 0.02, 0.12)
    s = MeshArrays().add_triangles(tris).scene()
    v, i, p = s.geometry()
    ref = O.RefScene(v, i, p)
    bvh = ref.bvh()
    o1, d1 = random_rays(2048, seed=203)
    # Pinhole rays from outside the cube towards it.
    g = np.random.default_rng(204)
    o2 = np.zeros((2048, 4), np.float32)
    d2 = np.zeros((2048, 4), np.float32)
    o2[:, :3] = (0.5, 0.5, 2.5)
    target = np.stack([g.uniform(0, 1, 2048), g.uniform(0, 1, 2048), np.zeros(2048)], axis=1)
    dd = target - o2[:, :3]
    d2[:, :3] = dd / np.linalg.norm(dd, axis=1, keepdims=True)
    o2[:, 3] = 1e-5
    d2[:, 3] = np.inf
    org = np.concatenate([o1, o2])
    dirs = np.concatenate([d1, d2])
    out = {"triangles": tris, "nodes": bvh.nodes.view(np.uint8).reshape(-1),
           "primitive_indices": bvh.primitive_indices, "max_depth": np.int32(bvh.max_depth),
           "org": org, "dir": dirs}
    out.update(trace_outputs(ref.trace(org, dirs)))
    return out


def two_level_fixture():
    s = Scene.instanced(instances=60, subdivisions=1, variants=3, env_width=8, env_height=4, seed=205)
    s.build_accel()
    v, i, p = s.geometry()
    _, _, inst = s.tlas()
    ip = inst["primitive"].copy()
    im = inst["object_to_world"].copy()
    ref = O.RefScene(v, i, p, ip, im)
    tn, tpi, _, _ = ref.tlas()
    org, dirs = s.camera_rays(80, 45)
    out = {"vertices": v.view(np.uint8).reshape(-1), "indices": i, "primitives": p.view(np.uint8).reshape(-1),
           "instance_primitive": ip, "instance_matrices": im, "tlas_nodes": tn.view(np.uint8).reshape(-1),
           "tlas_primitive_indices": tpi, "org": org, "dir": dirs}
    out.update(trace_outputs(ref.trace(org, dirs)))
    return out


def cornell_fixture():
    s = Scene.cornell(1)
    s.build_accel()
    film8, st8, _ = O.ref_render(s, 64, 64, 16, max_depth=8, sampler=2, seed=SEED)
    film1, st1, _ = O.ref_render(s, 64, 64, 4, max_depth=1, sampler=2, seed=SEED)
    keys = O.STAT_KEYS
    return {"digest": np.array(geometry_digest(s)), "width": np.int32(64), "height": np.int32(64),
            "film8": film8.view(np.uint8).reshape(-1), "stats8": np.array([st8[k] for k in keys], np.uint64),
            "spp8": np.int32(16), "film1": film1.view(np.uint8).reshape(-1),
            "stats1": np.array([st1[k] for k in keys], np.uint64), "spp1": np.int32(4),
            "seed": np.uint64(SEED)}


# Adaptive sampling settings of the fixture: (min_spp, relative_error,
# absolute_error, confidence) for depth 8 and depth 1.
ADAPTIVE8 = (4, 0.1, 0.001, 1.96)
ADAPTIVE1 = (4, 0.05, 0.001, 1.96)


def cornell_adaptive_fixture():
    s = Scene.cornell(1)
    s.build_accel()
    film8, st8, _ = O.ref_render(s, 64, 64, 32, max_depth=8, sampler=2, seed=SEED, adaptive=ADAPTIVE8)
    film1, st1, _ = O.ref_render(s, 64, 64, 16, max_depth=1, sampler=2, seed=SEED, adaptive=ADAPTIVE1)
    keys = O.STAT_KEYS
    return {"digest": np.array(geometry_digest(s)), "width": np.int32(64), "height": np.int32(64),
            "film8": film8.view(np.uint8).reshape(-1), "stats8": np.array([st8[k] for k in keys], np.uint64),
            "spp8": np.int32(32), "adaptive8": np.array(ADAPTIVE8, np.float32),
            "converged8": np.uint64(st8["converged_pixels"]),
            "film1": film1.view(np.uint8).reshape(-1), "stats1": np.array([st1[k] for k in keys], np.uint64),
            "spp1": np.int32(16), "adaptive1": np.array(ADAPTIVE1, np.float32),
            "converged1": np.uint64(st1["converged_pixels"]), "seed": np.uint64(SEED)}


def bsdf_fixture():
    """EvaluateBsdf / PdfBsdf / SampleBsdf / IsDeltaBsdf of the reference
    (src/shading/bsdf.cpp:17-329) on 256 random queries per material of
    tests/bsdf_cases.py (every in-scope branch) and on the known-answer cases."""
    from tests import bsdf_cases as B
    mi, q = B.queries(256, seed=12)
    cases = B.known_answer_cases()
    mats = list(B.MATERIALS) + [m for _, m, _ in cases]
    mi = np.concatenate([mi, np.arange(len(B.MATERIALS), len(mats), dtype=np.int32)])
    q = np.concatenate([q, np.stack([row for _, _, row in cases])])
    return {"kind": np.array([m["kind"] for m in mats], np.int32),
            "reflectance": np.array([m["reflectance"] for m in mats], np.float32),
            "ior": np.array([m.get("ior", 1.5) for m in mats], np.float32),
            "uroughness": np.array([m.get("uroughness", 0.0) for m in mats], np.float32),
            "vroughness": np.array([m.get("vroughness", m.get("uroughness", 0.0)) for m in mats], np.float32),
            "material_index": mi.astype(np.int32), "queries": q,
            "results": O.ref_bsdf_query(mats, mi, q)}


def zoo_film_fixture():
    """The reference's FP32 film + statistics of the material-zoo scene (every
    in-scope BSDF branch, sky environment): 64 x 36, 8 spp, depth 12."""
    from tests.helpers import material_zoo_scene
    s = material_zoo_scene()
    film, st, _ = O.ref_render(s, 64, 36, 8, max_depth=12, sampler=2, seed=SEED)
    keys = O.STAT_KEYS
    from tests.golden_data import scene_digest
    return {"digest": np.array(scene_digest(s)), "width": np.int32(64), "height": np.int32(36),
            "spp": np.int32(8), "depth": np.int32(12), "film": film.view(np.uint8).reshape(-1),
            "stats": np.array([st[k] for k in keys], np.uint64), "seed": np.uint64(SEED)}


def spheres_fixture():
    """Analytic spheres (yr::RenderSphere): the reference's per-ray hits (t,
    SphereUv, id = -2 - sphere) and node / triangle / sphere counters on the
    sphere Cornell box (flat BVH: spheres in leaves beside triangles) and the
    sphere zoo (spheres in the TLAS), and its FP32 film + statistics of the
    sphere Cornell box (48 x 36, 8 spp, depth 8)."""
    from tests.golden_data import scene_digest
    from tests.helpers import sphere_cornell, sphere_zoo
    out = {}
    c = sphere_cornell()
    v, i, p = c.geometry()
    sp = c.spheres()
    o1, d1 = c.camera_rays(64, 48)
    g = np.random.default_rng(206)
    o2 = np.zeros((2048, 4), np.float32)
    o2[:, :3] = g.uniform((-1, 0, -1), (1, 2, 1), (2048, 3))
    o2[:, 3] = 1e-5
    dd = g.normal(size=(2048, 3))
    d2 = np.zeros((2048, 4), np.float32)
    d2[:, :3] = dd / np.linalg.norm(dd, axis=1, keepdims=True)
    d2[:, 3] = np.inf
    org, dirs = np.concatenate([o1, o2]), np.concatenate([d1, d2])
    r = O.RefScene(v, i, p, spheres=sp).trace(org, dirs)
    out.update({"flat_" + k: val for k, val in trace_outputs(r).items()})
    out.update({"flat_sphere_tests": r["sphere_tests"], "flat_org": org, "flat_dir": dirs,
                "flat_digest": np.array(scene_digest(c))})
    film, st, _ = O.ref_render(c, 48, 36, 8, max_depth=8, sampler=2, seed=SEED)
    out.update({"film": film.view(np.uint8).reshape(-1), "stats": np.array([st[k] for k in O.STAT_KEYS], np.uint64),
                "width": np.int32(48), "height": np.int32(36), "spp": np.int32(8), "depth": np.int32(8),
                "seed": np.uint64(SEED)})
    z = sphere_zoo(60)
    v, i, p = z.geometry()
    _, _, inst = z.tlas()
    ref = O.RefScene(v, i, p, inst["primitive"].copy(), inst["object_to_world"].copy(), spheres=z.spheres())
    org, dirs = z.camera_rays(80, 45)
    r = ref.trace(org, dirs)
    out.update({"tl_" + k: val for k, val in trace_outputs(r).items()})
    out.update({"tl_sphere_tests": r["sphere_tests"], "tl_org": org, "tl_dir": dirs,
                "tl_digest": np.array(scene_digest(z))})
    return out


def read_yrir(path):
    """Sections of a `dropin_probe dump` file: name -> raw bytes (and count)."""
    with open(path, "rb") as f:
        data = f.read()
    assert data[:4] == b"YRIR"
    off, out = 4, {}
    while True:
        name = data[off:off + 16].split(b"\0")[0].decode()
        count = int(np.frombuffer(data, np.int64, 1, off + 16)[0])
        elem = int(np.frombuffer(data, np.int32, 1, off + 24)[0])
        off += 28
        if name == "end":
            return out
        out[name] = np.frombuffer(data, np.uint8, count * elem, off).copy()
        off += count * elem


def cornell_pbrt_fixture():
    """The reference's own Cornell scene (scenes/pbrt/cornell_box_pbrt: mirror
    and glass spheres, a ceiling area light and a point light) compiled by its
    PBRT frontend (the drop-in probe's `dump`), and the reference CPU backend's
    film + statistics of its 64 x 64 smoke variant at 32 spp, depth 6, rendered
    by the same probe."""
    import subprocess
    import tempfile

    from tests import dropin as D
    scenes = "/root/reference/scenes/pbrt/cornell_box_pbrt/"
    with tempfile.TemporaryDirectory() as tmp:
        out = {}
        for tag, name in (("full", "cornell_box_pbrt.pbrt"), ("smoke", "cornell_box_pbrt_smoke.pbrt")):
            path = os.path.join(tmp, tag + ".bin")
            subprocess.run([D.PROBE, "dump", scenes + name, path], check=True, cwd=tmp)
            ir = read_yrir(path)
            for k, v in ir.items():
                out[f"{tag}_{k}"] = v
        r, res = D.probe(["render", scenes + "cornell_box_pbrt_smoke.pbrt", "cpu", "32"],
                         os.path.join(tmp, "film.bin"), tmp)
        assert r.returncode == 0 and res["ok"], r.stderr
    out["film"] = res["film"].reshape(-1).view(np.uint8)
    out["stats"] = np.array([res["stats"][k] for k in O.STAT_KEYS], np.uint64)
    out["spp"] = np.int32(32)
    return out


def textures_pbrt_fixture():
    """Textured PBRT scenes compiled by the reference's frontend (the drop-in
    probe's `dump`) with the reference CPU backend's film + statistics at
    16 spp, rendered by the same probe: this repository's hip_textures.pbrt
    (EWA / trilinear / nearest / bilinear image textures, a normal map with
    tangents, textured thin glass, an HDR environment map) under "ours_", and
    the reference's own texture_test smoke scene (scenes/pbrt/texture_test:
    checker PNGs with repeat / clamp wraps, a normal-mapped plane without
    tangents, an HDR sky) under "tt_"."""
    import subprocess
    import tempfile

    from tests import dropin as D
    scenes = (("ours", os.path.join(D.SCENES, "hip_textures.pbrt")),
              ("tt", "/root/reference/scenes/pbrt/texture_test/texture_test_smoke.pbrt"))
    out = {}
    with tempfile.TemporaryDirectory() as tmp:
        for tag, path in scenes:
            ir_path = os.path.join(tmp, tag + ".bin")
            subprocess.run([D.PROBE, "dump", path, ir_path], check=True, cwd=tmp, capture_output=True)
            for k, v in read_yrir(ir_path).items():
                out[f"{tag}_{k}"] = v
            r, res = D.probe(["render", path, "cpu", "16"], os.path.join(tmp, tag + "_film.bin"), tmp)
            assert r.returncode == 0 and res["ok"], r.stderr
            out[f"{tag}_film"] = res["film"].reshape(-1).view(np.uint8)
            out[f"{tag}_stats"] = np.array([res["stats"][k] for k in O.STAT_KEYS], np.uint64)
    out["spp"] = np.int32(16)
    return out


FIXTURES = {"textures_pbrt": textures_pbrt_fixture, "cornell_pbrt": cornell_pbrt_fixture, "spheres": spheres_fixture, "sampler": sampler_fixture, "bsdf": bsdf_fixture, "zoo_film": zoo_film_fixture, "flat_trace": flat_fixture, "two_level_trace": two_level_fixture,
            "cornell_film": cornell_fixture, "cornell_adaptive": cornell_adaptive_fixture}


def main(names=None):
    if not O.ref_available():
        raise SystemExit("oracle/_ref/libyrref.so is missing: build it with `make -C oracle ref`")
    for name in names or FIXTURES:
        fn = FIXTURES[name]
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **fn())
        print(f"{path}: {os.path.getsize(path)} bytes")


if __name__ == "__main__":
    main(sys.argv[1:])
