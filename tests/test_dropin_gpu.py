"""The drop-in build on the GPU: the reference's own PBRT frontend, factory,
film and checkpoint code driving HipBackend.

* HIP vs the reference's CPU backend on the same PBRT scenes (every in-scope
  material kind as triangle meshes; analytic spheres from Shape "sphere";
  image textures with every filter, a normal map with tangents and an HDR
  environment map): image mean within 0.5 %, per-pixel RMSE
  below 0.35 x the CPU film's standard error, counters within 1 %, sample
  counts equal (same sampler streams; beyond the first bounce the device's
  transcendentals differ from glibc in the last bits).
* Checkpoint v3 round trip (src/film/film_checkpoint.cpp:71, 120): a HIP render
  cancelled after 5 passes, written with WriteFilmCheckpoint, loaded with
  LoadFilmCheckpoint and resumed, equals the uninterrupted render: sample
  counts on every pixel, sums within FP32 rounding (the loaded film is FP32,
  the device accumulates in FP64).
* The reference CLI renders the scene with --backend hip (its own CTest
  expectations, tests/cli_tests.cmake:66-70: "Rendered image:", "Hits: [^0]").
"""
import os
import re
import subprocess

import numpy as np
import pytest

from tests import dropin as D

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not D.built(), reason="drop-in build absent")]
LUM = np.array([0.2126, 0.7152, 0.0722])


def lum_mean(film):
    return (film["sum"].reshape(-1, 3) / np.maximum(film["samples"].reshape(-1), 1)[:, None]) @ LUM


@pytest.mark.parametrize("scene", ["hip_room.pbrt", "hip_spheres.pbrt", "hip_textures.pbrt"])
def test_hip_matches_reference_cpu_backend(tmp_path, scene):
    rc, cpu = D.render(scene, "cpu", tmp_path)
    rh, hip = D.render(scene, "hip", tmp_path)
    assert cpu["ok"] and hip["ok"], (rc.stderr, rh.stderr, hip and hip["error"])
    assert np.array_equal(hip["film"]["samples"], cpu["film"]["samples"])
    g, c = lum_mean(hip["film"]), lum_mean(cpu["film"])
    f = cpu["film"].reshape(-1)
    var = f["luminance_m2"] / np.maximum(f["samples"] - 1, 1)
    noise = float(np.sqrt(np.mean(np.maximum(var, 0) / np.maximum(f["samples"], 1))))
    rel_mean = abs(g.mean() - c.mean()) / c.mean()
    rmse = float(np.sqrt(np.mean((g - c) ** 2)))
    keys = ("rays_traced", "hits", "misses", "shadow_rays") + (("sphere_tests",) if "spheres" in scene else ())
    counts = {k: abs(hip["stats"][k] - cpu["stats"][k]) / max(1, cpu["stats"][k]) for k in keys}
    print(f"drop-in {scene}: rel mean {rel_mean:.5f} rmse {rmse:.5f} noise {noise:.5f} counts {counts}", flush=True)
    assert rel_mean < 0.005 and rmse < 0.35 * noise and max(counts.values()) < 0.01


def test_checkpoint_round_trip(tmp_path):
    rr, full = D.render("hip_room.pbrt", "hip", tmp_path)
    rk, resumed, ckpt = D.checkpoint("hip_room.pbrt", "hip", tmp_path, 5)
    assert full["ok"] and rk.returncode == 0 and resumed["ok"], (rk.stdout, rk.stderr)
    assert os.path.getsize(ckpt) > 0
    a, b = full["film"].reshape(-1), resumed["film"].reshape(-1)
    assert np.array_equal(a["samples"], b["samples"])
    assert np.allclose(a["sum"], b["sum"], rtol=4 * 2.0 ** -24, atol=1e-30)
    assert np.array_equal(a["aov_samples"], b["aov_samples"])


def test_reference_cli_renders_with_hip(tmp_path):
    r = subprocess.run([D.CLI, "render", os.path.join(D.SCENES, "hip_room.pbrt"), "--backend", "hip"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=600)
    print(r.stdout[-600:], flush=True)
    assert r.returncode == 0, r.stderr
    assert "Integrator: path" in r.stdout and "Rendered image:" in r.stdout
    assert re.search(r"Hits: [1-9]", r.stdout)
