"""Analytic (point / spot / distant) lights on the GPU against the oracle
(pinned to the reference's SampleAnalyticLight and next-event term bit for bit
in test_analytic_lights.py), through the C ABI.

  * depth-1 renders (camera hit + next-event estimation of the emissive,
    analytic and environment lights; light selection by alias table or light
    tree uses + - * / only): ray, hit, shadow and occlusion counters, sample
    counts and depth AOVs equal, radiance within 1e-5 relative (the
    conductor's Schlick term calls pow); traversal counters equal where shadow
    rays walk closest hits (scenes with glass), not above the oracle's where
    they stop at their first hit (opaque scenes);
  * full renders: statistical parity (image mean within 0.5 %, per-pixel RMSE
    below 0.35 x the oracle's per-pixel standard error, counters within 1 %).
"""
import numpy as np
import pytest

from oracle import oracle_py as O
from tests.helpers import cornell_lights_scene, lit_sphere_scene, many_point_lights_scene
from tests.test_render_gpu import gpu_render, statistical_parity

pytestmark = pytest.mark.gpu

SCENES = {
    "cornell_spheres": lambda: cornell_lights_scene(True),
    "tree": lambda: many_point_lights_scene(12),
    "lit_sphere": lambda: lit_sphere_scene([{"kind": "distant", "direction": (0.2, -1.0, -0.3),
                                             "intensity": (5.0, 5.0, 5.0)},
                                            {"kind": "point", "position": (0.0, 2.0, 1.0),
                                             "intensity": (40.0, 40.0, 40.0)}]),
}


@pytest.mark.parametrize("kind", sorted(SCENES))
def test_depth1_analytic_render_matches_oracle(device, kind):
    s = SCENES[kind]()
    print(f"depth-1 analytic render {kind} 64x48x2", flush=True)
    g, gs = gpu_render(device, s, 64, 48, 2, max_depth=1)
    o, os_ = O.render(s, 64, 48, 2, max_depth=1)
    for k in ("rays_traced", "hits", "misses", "shadow_rays", "occluded_shadow_rays", "bad_samples"):
        assert gs[k] == os_[k], k
    for k in ("triangle_tests", "sphere_tests", "bvh_node_tests"):
        if kind == "cornell_spheres":  # glass: shadow rays walk closest hits, as the reference's do
            assert gs[k] == os_[k], k
        else:  # opaque: shadow rays stop at their first hit (same visibility, fewer tests)
            assert gs[k] <= os_[k], k
    assert np.array_equal(g["samples"], o["samples"]) and np.array_equal(g["aov_samples"], o["aov_samples"])
    assert np.array_equal(g["depth_sum"].astype(np.float32), o["depth_sum"])
    assert np.allclose(g["sum"], o["sum"], rtol=1e-5, atol=1e-7)
    assert os_["shadow_rays"] > 0 and np.any(o["sum"] > 0)


@pytest.mark.parametrize("kind", ["cornell_spheres", "tree"])
def test_analytic_render_statistical_parity(device, kind):
    s = SCENES[kind]()
    print(f"analytic render {kind} 64x48x16 depth 8", flush=True)
    rel_mean, rmse, noise, counts = statistical_parity(device, s, 64, 48, 16, 8)
    print(f"analytic {kind}: rel mean {rel_mean:.5f} rmse {rmse:.5f} noise {noise:.5f} counts {counts}", flush=True)
    assert rel_mean < 0.005
    assert rmse < 0.35 * noise
    assert max(counts.values()) < 0.01
