"""GPU parity of the wavefront path tracer against the CPU oracle (pinned
bit-for-bit to the reference's RenderCpuPathTrace in test_oracle_render.py).

Exact where the arithmetic is exact: sampler values, camera rays, and whole
max-depth-1 renders of the Cornell workload (emission + emissive NEE use only
+ - * / sqrt). Beyond the first bounce the device's cos/sin/pow/atan2 differ
from glibc in the last bits, so directions drift and parity is statistical:
  * image mean luminance within 0.5 %,
  * per-pixel RMSE(GPU - oracle) below 0.35 x the oracle's per-pixel standard
    error (same sample streams, so most paths coincide),
  * counters (rays, hits, misses, shadow rays) within 1 %.
"""
import numpy as np
import pytest

from oracle import oracle_py as O
from tests.helpers import base_path_scene
from yaoray_amd import Scene

pytestmark = pytest.mark.gpu


def gpu_render(device, scene, w, h, spp, **kw):
    film, stats = device.render(scene, w, h, spp, **kw)
    return film, stats


def mean_image(film, w, h):
    return (film["sum"] / np.maximum(film["samples"], 1)[:, None]).reshape(h, w, 3)


@pytest.mark.parametrize("kind", [0, 1, 2, 3])
def test_sampler_values_bit_exact(device, kind):
    rng = np.random.default_rng(kind)
    n = 20000
    x, y = rng.integers(0, 4096, n), rng.integers(0, 4096, n)
    sample, dim = rng.integers(0, 1024, n), rng.integers(0, 200, n)
    for spp in (1, 7, 16):
        for dims in (1, 2):
            g = device.sample_values(kind, 0x59414F524159, spp, x, y, sample, dim, dims)
            o = O.sample_values(kind, 0x59414F524159, spp, x, y, sample, dim, dims)
            assert np.array_equal(g.view(np.uint32), o.view(np.uint32)), (spp, dims)


def test_camera_rays_bit_exact(device):
    s = Scene.cornell(1)
    s.build_accel()
    for spp, sample in ((1, 0), (16, 5)):
        rd = s.render_desc(96, 64, spp, sample_index=sample)
        go, gd = device.camera_rays(rd)
        oo, od = O.camera_rays(rd)
        assert np.array_equal(go.view(np.uint32), oo.view(np.uint32))
        assert np.array_equal(gd.view(np.uint32), od.view(np.uint32))


def test_depth1_cornell_bit_exact(device):
    s = Scene.cornell(1)
    s.build_accel()
    g, gs = gpu_render(device, s, 64, 64, 1, max_depth=1)
    o, os_ = O.render(s, 64, 64, 1, max_depth=1)
    assert np.array_equal(g["sum"].astype(np.float32).view(np.uint32), o["sum"].view(np.uint32))
    assert np.array_equal(g["samples"], o["samples"])
    for k in ("rays_traced", "hits", "misses", "shadow_rays", "occluded_shadow_rays", "bad_samples"):
        assert gs[k] == os_[k], k
    assert np.array_equal(g["aov_samples"], o["aov_samples"])
    assert np.array_equal(g["depth_sum"].astype(np.float32), o["depth_sum"])


def test_known_answers(device):
    s = base_path_scene()
    s.set_materials([{"kind": 0, "reflectance": (0.0, 0.0, 0.0), "emission": (0.25, 0.5, 0.75)}])
    s.set_constant_environment((0.0, 0.0, 0.0))
    film, _ = gpu_render(device, s, 3, 3, 1, max_depth=1, sampler=0, seed=7)
    assert np.allclose(mean_image(film, 3, 3)[1, 1], [0.25, 0.5, 0.75], atol=1e-6)
    s = base_path_scene()
    film, st = gpu_render(device, s, 4, 3, 1, max_depth=1, sampler=0, seed=7)
    assert st["rays_traced"] == 12 and st["hits"] + st["misses"] == 12
    film, st = gpu_render(device, s, 2, 2, 4, max_depth=1, sampler=0, seed=7)
    assert st["rays_traced"] == 16 and np.all(film["samples"] == 4)
    env = np.array([0.02, 0.03, 0.04])
    film, _ = gpu_render(device, s, 9, 9, 4, max_depth=2, sampler=2, seed=3)
    img = mean_image(film, 9, 9)
    assert np.allclose(img[0, 0], env, rtol=1e-6) and np.allclose(img[4, 4], 0.8 * env, rtol=1e-5)


def statistical_parity(device, scene, w, h, spp, depth, sampler=2):
    g, gs = gpu_render(device, scene, w, h, spp, max_depth=depth, sampler=sampler)
    o, os_ = O.render(scene, w, h, spp, max_depth=depth, sampler=sampler)
    gi, oi = mean_image(g, w, h), mean_image(o, w, h)
    lum = lambda im: im @ np.array([0.2126, 0.7152, 0.0722])
    gl, ol = lum(gi), lum(oi)
    rel_mean = abs(gl.mean() - ol.mean()) / ol.mean()
    var = o["luminance_m2"] / np.maximum(o["samples"] - 1, 1)
    stderr = np.sqrt(np.maximum(var, 0) / np.maximum(o["samples"], 1)).reshape(h, w)
    rmse = np.sqrt(np.mean((gl - ol) ** 2))
    noise = np.sqrt(np.mean(stderr ** 2))
    counts = {k: abs(gs[k] - os_[k]) / max(1, os_[k]) for k in ("rays_traced", "hits", "misses", "shadow_rays")}
    return rel_mean, rmse, noise, counts


def test_cornell_statistical_parity(device):
    s = Scene.cornell(1)
    s.build_accel()
    rel_mean, rmse, noise, counts = statistical_parity(device, s, 64, 64, 16, 8)
    print(f"cornell: rel mean {rel_mean:.5f} rmse {rmse:.5f} noise {noise:.5f} counts {counts}")
    assert rel_mean < 0.005
    assert rmse < 0.35 * noise
    assert max(counts.values()) < 0.01


def test_two_level_environment_statistical_parity(device):
    s = Scene.instanced(instances=300, subdivisions=3, variants=6, env_width=256, env_height=128, seed=4)
    s.build_accel()
    rel_mean, rmse, noise, counts = statistical_parity(device, s, 96, 54, 8, 12)
    print(f"two-level: rel mean {rel_mean:.5f} rmse {rmse:.5f} noise {noise:.5f} counts {counts}")
    assert rel_mean < 0.005
    assert rmse < 0.35 * noise
    assert max(counts.values()) < 0.01


@pytest.mark.parametrize("kind", ["cornell", "two_level"])
def test_multi_sample_pass_equals_single_passes(device, kind):
    """A pass carrying several sample indices leaves the same FP64 film and counters
    as one pass per sample index."""
    if kind == "cornell":
        s = Scene.cornell(1)
        w, h, spp, depth = 40, 30, 7, 8
    else:
        s = Scene.instanced(instances=40, subdivisions=2, variants=3, env_width=64, env_height=32, seed=5)
        w, h, spp, depth = 32, 20, 5, 10
    s.build_accel()
    one, st1 = device.render(s, w, h, spp, max_depth=depth)
    many, st3 = device.render(s, w, h, spp, max_depth=depth, batch=3)
    for field in one.dtype.names:
        assert np.array_equal(one[field], many[field]), field
    assert st1 == st3
