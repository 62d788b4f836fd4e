"""Textures, alpha cut-outs and normal maps on the GPU against the oracle
(pinned to the reference's SampleTexture4, mip chains and textured, cut-out
and normal-mapped renders bit for bit in test_textures.py), through the C ABI.

  * yrh_texture_query: every filter and wrap mode at random points and
    footprints; nearest and bilinear lookups bit-exact, trilinear and EWA
    within 2e-6 (their level-of-detail weights use hypot and log2, which the
    device evaluates in double and the host's libm in float: the last bit of
    a weight may differ) and bit-exact on at least 99 % of the points;
  * depth-1 renders (camera hits with EWA / trilinear footprints, alpha skips
    in the camera rays and in the shadow walk, normal maps, textured thin
    glass in shadows): every counter, sample count and depth AOV equal,
    radiance within 2e-5 relative;
  * full renders: statistical parity (image mean within 0.5 %, per-pixel RMSE
    below 0.35 x the oracle's standard error, counters within 1 %), and a
    guided render's mean equal to the plain render's (paired t test).
"""
import numpy as np
import pytest

from oracle import oracle_py as O
from tests.helpers import alpha_stack_scene, guide_desc, textured_scene
from tests.test_render_gpu import gpu_render, statistical_parity
from yaoray_amd import native as N

pytestmark = pytest.mark.gpu

FILTERS = (N.FILTER_NEAREST, N.FILTER_BILINEAR, N.FILTER_TRILINEAR, N.FILTER_EWA)
WRAPS = (N.WRAP_REPEAT, N.WRAP_CLAMP, N.WRAP_MIRRORED)


def random_queries(n, seed, zero=0):
    rng = np.random.default_rng(seed)
    uv = (rng.random((n, 2), dtype=np.float32) * 6.0 - 3.0).astype(np.float32)
    scale = (10.0 ** rng.uniform(-4.0, 0.5, (n, 1))).astype(np.float32)
    fp = (rng.standard_normal((n, 4)).astype(np.float32) * scale).astype(np.float32)
    fp[:zero] = 0.0
    return uv, fp


@pytest.mark.parametrize("shape", [(1, 1), (2, 2), (3, 5), (16, 9), (64, 64)])
def test_texture_query_matches_oracle(device, shape):
    tex = np.random.default_rng(shape[0] * 100 + shape[1]).random((*shape, 4), dtype=np.float32)
    print(f"texture query {shape}", flush=True)
    for ws in WRAPS:
        for wt in WRAPS:
            for f in FILTERS:
                uv, fp = random_queries(2000, 7 * ws + wt + 31 * f, zero=100)
                g = device.texture_query(tex, uv, fp, f, ws, wt)
                o = O.sample_texture(tex, uv, fp, f, ws, wt)
                same = (g.view(np.uint32) == o.view(np.uint32)).all(axis=1)
                if f in (N.FILTER_NEAREST, N.FILTER_BILINEAR):
                    assert same.all(), (ws, wt, f)
                else:
                    assert np.allclose(g, o, rtol=2e-6, atol=2e-6), (ws, wt, f, np.abs(g - o).max())
                    assert same.mean() >= 0.99, (ws, wt, f, same.mean())
    g = device.texture_query(tex, uv, None, N.FILTER_EWA)
    assert np.array_equal(g.view(np.uint32), O.sample_texture(tex, uv, None, N.FILTER_EWA).view(np.uint32))


def test_empty_texture_samples_zero(device):
    uv, fp = random_queries(16, 3)
    g = device.texture_query(np.zeros((0, 0, 4), np.float32), uv, fp)
    assert np.array_equal(g, np.tile(np.array([0, 0, 0, 1], np.float32), (16, 1)))


SCENES = {
    "flat": lambda: textured_scene(),
    "two_level": lambda: textured_scene(two_level=True),
    "bilinear_only": lambda: textured_scene(filters=(N.FILTER_BILINEAR,) * 4),
    "stack70": lambda: alpha_stack_scene(layers=70),
    "stack30": lambda: alpha_stack_scene(layers=30),
}


@pytest.mark.parametrize("kind", sorted(SCENES))
def test_depth1_textured_render_matches_oracle(device, kind):
    s = SCENES[kind]()
    print(f"depth-1 textured render {kind} 64x48x2", flush=True)
    g, gs = gpu_render(device, s, 64, 48, 2, max_depth=1)
    o, os_ = O.render(s, 64, 48, 2, max_depth=1)
    for k in ("rays_traced", "hits", "misses", "shadow_rays", "occluded_shadow_rays", "bad_samples", "triangle_tests",
              "sphere_tests", "bvh_node_tests"):
        assert gs[k] == os_[k], (k, gs[k], os_[k])
    assert np.array_equal(g["samples"], o["samples"]) and np.array_equal(g["aov_samples"], o["aov_samples"])
    assert np.array_equal(g["depth_sum"].astype(np.float32), o["depth_sum"])
    assert np.allclose(g["sum"], o["sum"], rtol=2e-5, atol=1e-7)
    assert np.allclose(g["albedo_sum"], o["albedo_sum"], rtol=2e-5, atol=1e-7)
    assert np.allclose(g["normal_sum"], o["normal_sum"], rtol=2e-5, atol=1e-6)
    assert os_["shadow_rays"] > 0 and np.any(o["sum"] > 0)


@pytest.mark.parametrize("kind", ["flat", "two_level", "stack30"])
def test_textured_render_statistical_parity(device, kind):
    s = SCENES[kind]()
    print(f"textured render {kind} 64x48x16 depth 6", flush=True)
    rel_mean, rmse, noise, counts = statistical_parity(device, s, 64, 48, 16, 6)
    print(f"textured {kind}: rel mean {rel_mean:.5f} rmse {rmse:.5f} noise {noise:.5f} counts {counts}", flush=True)
    assert rel_mean < 0.005
    assert rmse < 0.35 * noise
    assert max(counts.values()) < 0.01


def test_guided_textured_render_matches_plain_mean(device):
    """The guide's queries resolve textured, normal-mapped surfaces too; the
    guided estimator keeps the plain one's expectation (paired differences
    over 8 sampler seeds, two-sided 0.1 % t bound for 7 degrees of freedom)."""
    from tests.test_guide_gpu import configure

    s = textured_scene()
    w, h, spp = 64, 48, 16
    configure(device, guide_desc(log2_table=14, rmin=4, rmax=64, bounds=((-4.5, -0.5, -6.5), (4.5, 3.5, 3.5))))
    device.upload_scene(s)
    device.upload_shading(s)
    device.guide_set_mode(N.GUIDE_RECORD)
    try:
        for k in range(3):
            device.render(s, w, h, 4, max_depth=6, upload=False, seed=700 + k)
            device.guide_train(16, 8192)
    finally:
        device.guide_set_mode(N.GUIDE_OFF)
    lum = np.array([0.2126, 0.7152, 0.0722])
    diffs, means = [], []
    for seed in range(8):
        plain, _ = device.render(s, w, h, spp, max_depth=6, upload=False, seed=seed)
        device.guide_set_mode(N.GUIDE_SAMPLE)
        try:
            guided, _ = device.render(s, w, h, spp, max_depth=6, upload=False, seed=seed)
        finally:
            device.guide_set_mode(N.GUIDE_OFF)
        gl = ((guided["sum"] / np.maximum(guided["samples"], 1)[:, None]) @ lum).mean()
        pl = ((plain["sum"] / np.maximum(plain["samples"], 1)[:, None]) @ lum).mean()
        diffs.append(gl - pl)
        means.append(pl)
    diffs = np.array(diffs)
    t = diffs.mean() / max(diffs.std(ddof=1) / np.sqrt(len(diffs)), 1e-12)
    print(f"textured guided - plain: {diffs.mean():.5f} (plain {np.mean(means):.5f}), t = {t:.2f}", flush=True)
    assert abs(t) < 5.408  # two-sided 0.1 % for 7 degrees of freedom
    assert abs(diffs.mean()) < 0.02 * np.mean(means)
