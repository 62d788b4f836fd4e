"""Loaders of the golden fixtures in tests/golden/ (written from the reference
by tests/golden/make_golden.py): inputs rebuilt into scenes, outputs as arrays."""
import hashlib
import os

import numpy as np

from tests.helpers import MeshArrays
from yaoray_amd import Scene
from yaoray_amd.api import FILM32_DTYPE, NODE_DTYPE, PRIMITIVE_DTYPE, VERTEX_DTYPE

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
STAT_KEYS = ("rays_traced", "shadow_rays", "occluded_shadow_rays", "triangle_tests", "sphere_tests",
             "bvh_node_tests", "hits", "misses", "bad_samples")


def load(name):
    with np.load(os.path.join(HERE, name + ".npz"), allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def expected_hits(f, prefix=""):
    h = f[prefix + "hit"]
    out = {"t": h[:, 0].copy(), "u": h[:, 1].copy(), "v": h[:, 2].copy(),
           "triangle": h[:, 3].copy().view(np.int32), "instance": f[prefix + "instance"],
           "node_tests": f[prefix + "node_tests"], "triangle_tests": f[prefix + "triangle_tests"]}
    if prefix + "sphere_tests" in f:
        out["sphere_tests"] = f[prefix + "sphere_tests"]
    return out


def flat_scene(f):
    s = MeshArrays().add_triangles(f["triangles"]).scene()
    s.build_accel()
    return s


def two_level_scene(f):
    v = f["vertices"].view(VERTEX_DTYPE)
    p = f["primitives"].view(PRIMITIVE_DTYPE)
    s = Scene.from_arrays(v, f["indices"], p, f["instance_primitive"], f["instance_matrices"])
    s.build_accel()
    return s


def nodes(raw):
    return raw.view(NODE_DTYPE)


def film(raw):
    return raw.view(FILM32_DTYPE)


def stats(arr):
    return dict(zip(STAT_KEYS, (int(x) for x in arr)))


def geometry_digest(scene):
    v, i, p = scene.geometry()
    h = hashlib.sha256()
    for a in (v, i, p):
        h.update(np.ascontiguousarray(a).tobytes())
    h.update(np.array(list(scene.info.camera), np.float32).tobytes())
    return h.hexdigest()


def cornell_scene(f):
    s = Scene.cornell(1)
    s.build_accel()
    assert geometry_digest(s) == str(f["digest"]), "Cornell workload inputs differ from the fixture's"
    return s


def bsdf_materials(f):
    """Material dicts of the BSDF fixture (kind, reflectance, ior, roughness)."""
    return [{"kind": int(k), "reflectance": tuple(float(x) for x in r), "ior": float(i), "uroughness": float(u),
             "vroughness": float(v)}
            for k, r, i, u, v in zip(f["kind"], f["reflectance"], f["ior"], f["uroughness"], f["vroughness"])]


def scene_digest(scene):
    """Digest of geometry, camera, instances, spheres and materials (the key of the
    fixtures of procedural scenes)."""
    import ctypes

    from yaoray_amd import native as N
    h = hashlib.sha256(geometry_digest(scene).encode())
    if scene.info.accel_kind == 1:
        _, _, inst = scene.tlas()
        h.update(np.ascontiguousarray(inst).tobytes())
    sph = scene.spheres()
    if len(sph):
        h.update(sph.tobytes())
    sd = scene.shading_desc()
    size = ctypes.sizeof(N.Material)
    raw = ctypes.string_at(sd.materials, sd.material_count * size)
    untextured = np.array([-1, -1, -1, np.float32(1.0).view(np.int32)], np.int32).tobytes()
    for k in range(sd.material_count):
        rec = raw[k * size:(k + 1) * size]
        # the texture fields enter the digest only when a material has textures
        h.update(rec[:64] if rec[64:] == untextured else rec)
    return h.hexdigest()


def zoo_scene(f):
    from tests.helpers import material_zoo_scene
    s = material_zoo_scene()
    assert scene_digest(s) == str(f["digest"]), "material-zoo inputs differ from the fixture's"
    return s


def sphere_scenes(f):
    """The sphere Cornell box and the sphere zoo of the spheres fixture."""
    from tests.helpers import sphere_cornell, sphere_zoo
    c, z = sphere_cornell(), sphere_zoo(60)
    assert scene_digest(c) == str(f["flat_digest"]), "sphere Cornell inputs differ from the fixture's"
    assert scene_digest(z) == str(f["tl_digest"]), "sphere zoo inputs differ from the fixture's"
    return c, z


def pbrt_scene(f, tag="smoke"):
    """A Scene from a PBRT-compiled RenderSceneIR fixture (the drop-in probe's
    dump); returns (scene, settings dict)."""
    prefix = tag + "_"
    return Scene.from_ir({k[len(prefix):]: v for k, v in f.items() if k.startswith(prefix)})
