"""Analytic spheres on the GPU, through the C ABI, against the oracle (pinned
to the reference's IntersectSphere / TLAS sphere slots / sphere surface path
bit for bit in test_spheres.py).

  * ray casts: hit ids, t, instance and node / triangle / sphere counters bit
    for bit; barycentrics of triangle hits bit for bit; a sphere hit's (u, v)
    (SphereUv: atan2 and acos of the same bits, computed by the device's math
    library) within 4e-6 (a few ulp of values in [0, 1]);
  * any-hit: occluded == (closest hit found), spheres included;
  * depth-1 renders (camera hit + next-event estimation, no transcendental
    except in rough-conductor terms): counters, sample counts and depth AOVs
    bit for bit, radiance within 1e-5 relative;
  * full renders: statistical parity (image mean within 0.5 %, per-pixel RMSE
    below 0.35 x the oracle's per-pixel standard error, counters within 1 %).
"""
import numpy as np
import pytest

from oracle import oracle_py as O
from tests.helpers import same_bits, sphere_cornell, sphere_zoo
from tests.test_render_gpu import gpu_render, statistical_parity

pytestmark = pytest.mark.gpu


def cast_rays(scene, kind, n=30000, seed=5):
    rng = np.random.default_rng(seed)
    o1, d1 = scene.camera_rays(160, 120)
    lo, hi = ((-1, 0, -1), (1, 2, 1)) if kind == "flat" else ((-12, 0, 5), (12, 6, 40))
    o2 = np.zeros((n, 4), np.float32)
    o2[:, :3] = rng.uniform(lo, hi, (n, 3))
    o2[:, 3] = 1e-5
    dd = rng.normal(size=(n, 3))
    d2 = np.zeros((n, 4), np.float32)
    d2[:, :3] = dd / np.linalg.norm(dd, axis=1, keepdims=True)
    d2[:, 3] = np.inf
    return np.concatenate([o1, o2]), np.concatenate([d1, d2])


def oracle_cast(scene, kind, org, d, occlusion=False):
    v, i, p = scene.geometry()
    sp = scene.spheres()
    if kind == "flat":
        return O.trace_flat(v, i, p, scene.flat_bvh(), org, d, occlusion=occlusion, spheres=sp)
    tl, tp, inst = scene.tlas()
    return O.trace_two_level(v, i, p, [scene.blas(k) for k in range(scene.info.blas_count)], inst, tl, tp, org, d,
                             occlusion=occlusion, spheres=sp)


def assert_hits(g, o, label):
    for k in ("t", "triangle", "node_tests", "triangle_tests", "sphere_tests", "instance"):
        if k in o:
            assert same_bits(g[k], o[k]), (label, k, int(np.sum(g[k] != o[k])))
    tri = o["triangle"] >= 0
    sph = o["triangle"] <= -2
    for k in ("u", "v"):
        assert same_bits(g[k][tri], o[k][tri]), (label, k)
        err = np.abs(g[k][sph].astype(np.float64) - o[k][sph])
        print(f"{label} sphere {k}: max |diff| {err.max():.2e}, bit-equal {np.mean(g[k][sph] == o[k][sph]):.4f}",
              flush=True)
        assert err.max() <= 4e-6, (label, k)
    assert sph.sum() > 1000 and tri.sum() > 1000


@pytest.mark.parametrize("kind", ["flat", "two_level"])
def test_sphere_casts_match_oracle(device, kind):
    s = sphere_cornell() if kind == "flat" else sphere_zoo(150)
    org, d = cast_rays(s, kind)
    print(f"sphere casts {kind}: {len(org)} rays", flush=True)
    device.upload_scene(s)
    g = device.trace_closest(org, d, counters=True)
    o = oracle_cast(s, kind, org, d)
    assert_hits(g, o, kind)
    ga = device.trace_any(org, d, counters=True)
    oa = oracle_cast(s, kind, org, d, occlusion=True)
    assert np.array_equal(ga["occluded"], oa["occluded"])
    assert np.array_equal(ga["occluded"].astype(bool), o["triangle"] != -1)
    assert same_bits(ga["node_tests"], oa["node_tests"]) and same_bits(ga["triangle_tests"], oa["triangle_tests"])


def test_depth1_sphere_render_matches_oracle(device):
    s = sphere_cornell()
    print("depth-1 sphere render 64x48x2", flush=True)
    g, gs = gpu_render(device, s, 64, 48, 2, max_depth=1)
    o, os_ = O.render(s, 64, 48, 2, max_depth=1)
    for k in ("rays_traced", "hits", "misses", "shadow_rays", "occluded_shadow_rays", "bad_samples",
              "triangle_tests", "sphere_tests", "bvh_node_tests"):
        assert gs[k] == os_[k], k
    assert np.array_equal(g["samples"], o["samples"]) and np.array_equal(g["aov_samples"], o["aov_samples"])
    assert np.array_equal(g["depth_sum"].astype(np.float32), o["depth_sum"])
    assert np.allclose(g["sum"], o["sum"], rtol=1e-5, atol=1e-7)
    assert os_["sphere_tests"] > 0


@pytest.mark.parametrize("kind", ["flat", "two_level"])
def test_sphere_render_statistical_parity(device, kind):
    if kind == "flat":
        s, w, h, spp, depth = sphere_cornell(), 64, 48, 16, 8
    else:
        s, w, h, spp, depth = sphere_zoo(150), 96, 54, 8, 12
    print(f"sphere render {kind} {w}x{h}x{spp} depth {depth}", flush=True)
    rel_mean, rmse, noise, counts = statistical_parity(device, s, w, h, spp, depth)
    print(f"spheres {kind}: rel mean {rel_mean:.5f} rmse {rmse:.5f} noise {noise:.5f} counts {counts}", flush=True)
    assert rel_mean < 0.005
    assert rmse < 0.35 * noise
    assert max(counts.values()) < 0.01
