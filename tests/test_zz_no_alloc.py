"""yrh_trace_closest_noalloc, the allocation-free closest-hit entry point,
against yrh_trace_closest through the C ABI: hits (t, u, v, triangle index),
instance indices and node / triangle counters bit for bit, on two calls in a
row (the second a random subset of the first's rays, so it reuses the
scratch array the first left behind), for a flat tree, a two-level scene, a two-level scene with analytic
spheres in its TLAS and a tree deep enough to spill its traversal stack into
the scratch array."""
import ctypes as C

import numpy as np
import pytest

from tests.helpers import MeshArrays, rays, sphere_zoo
from yaoray_amd import Scene, random_rays
from yaoray_amd import native as N

pytestmark = pytest.mark.gpu


def trace(device, fn_name, org, dirs):
    n = len(org)
    d_org = device.to_device(np.ascontiguousarray(org, np.float32))
    d_dir = device.to_device(np.ascontiguousarray(dirs, np.float32))
    d_hit, d_inst, d_nt, d_tt = device.alloc(16 * n), device.alloc(4 * n), device.alloc(4 * n), device.alloc(4 * n)
    soa = N.RaySoa(d_org.ptr, d_dir.ptr)
    hits = N.HitSoa(d_hit.ptr, d_inst.ptr, d_nt.ptr, d_tt.ptr)
    fn = getattr(N.yrh(), fn_name)
    rc = fn(device.ctx, C.byref(soa), n, C.byref(hits))
    assert rc == 0, (fn_name, N.yrh().yrh_last_error(device.ctx))
    device.synchronize()
    return {"hit": d_hit.download(np.float32, (n, 4)).view(np.uint32), "instance": d_inst.download(np.int32, (n,)),
            "node_tests": d_nt.download(np.uint32, (n,)), "triangle_tests": d_tt.download(np.uint32, (n,))}


def check_two_calls(device, scene, org, dirs, n_small):
    device.upload_scene(scene)
    sub = np.sort(np.random.default_rng(7).choice(len(org), n_small, replace=False))
    for k, (o, d) in enumerate(((org, dirs), (org[sub], dirs[sub]))):
        print(f"no-alloc call {k}: {len(o)} rays", flush=True)
        a = trace(device, "yrh_trace_closest_noalloc", o, d)
        b = trace(device, "yrh_trace_closest", o, d)
        for key in a:
            assert np.array_equal(a[key], b[key]), (k, key)
        assert (a["hit"][:, 3].view(np.int32) >= 0).any()


def test_noalloc_flat(device):
    s = Scene.triangles(200000, seed=2)
    s.build_accel()
    o1, d1 = s.camera_rays(256, 256)
    o2, d2 = random_rays(200000, seed=3)
    org, dirs = np.concatenate([o1, o2]), np.concatenate([d1, d2])
    perm = np.random.default_rng(1).permutation(len(org))
    check_two_calls(device, s, org[perm], dirs[perm], 5000)


def test_noalloc_two_level(device):
    s = Scene.instanced(instances=500, subdivisions=3, variants=4, env_width=8, env_height=4, seed=4)
    s.build_accel()
    o1, d1 = s.camera_rays(320, 180)
    check_two_calls(device, s, o1, d1, 777)


def test_noalloc_spheres(device):
    s = sphere_zoo(150)
    o1, d1 = s.camera_rays(200, 120)
    check_two_calls(device, s, o1, d1, 1500)


def test_noalloc_deep_tree_spills_into_scratch(device):
    tris = []
    for k in range(59):
        x = 4.0 ** k
        tris.append([[x, 0, 0], [x * 1.05, 0, 0], [x, 0.5 * max(1.0, x * 0.01), 0]])
    s = MeshArrays().add_triangles(np.array(tris, np.float32)).scene()
    s.build_accel()
    assert s.flat_bvh().max_depth > 20
    xs = np.array([4.0 ** k * 1.01 for k in range(59)] * 4, np.float32)
    ys = np.repeat([0.05, 0.1, 0.2, 0.45], 59).astype(np.float32)
    org, d = rays(np.stack([xs, ys, np.full_like(xs, 1.0)], 1), [(0, 0, -1)] * len(xs))
    o2, d2 = rays([(-1.0, 0.1, 0.0)] * 4, [(1, 0, 0), (1, 0.001, 0), (1, 0, 0.001), (1, 0.01, 0)])
    check_two_calls(device, s, np.concatenate([o2, org]), np.concatenate([d2, d]), 40)
