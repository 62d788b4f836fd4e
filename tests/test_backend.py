"""yr::HipBackend behind the reference's RenderBackend boundary
(src/backends/cpu/cpu_backend.cpp:54-102, cpu_path_tracer.cpp:39-229).

CPU tests cover the paths that end before any GPU work: a job requested for
another backend, and no HIP device. The GPU tests check that Render drives
the same kernels as the pass-level C ABI (same film, bit for bit), and mirror
the reference's progress / cancel / resume tests
(tests/backends/cpu/cpu_path_tracer_test.cpp, tests/runtime/*).
"""
import numpy as np
import pytest

from yaoray_amd import Scene
from yaoray_amd.api import FILM32_DTYPE


def cornell():
    s = Scene.cornell(1)
    return s


def test_prepare_rejects_other_backend():
    rep, film = cornell().backend_render(8, 8, 2, requested_backend=1)
    assert rep["prepare_ok"] == 0 and rep["ok"] == 0 and film is None
    assert rep["error"] == "HIP backend cannot prepare a scene requested for another backend"


def test_prepare_without_device_fails_loudly():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    rep, film = cornell().backend_render(8, 8, 2)
    assert rep["prepare_ok"] == 0 and film is None
    assert rep["error"].startswith("no HIP device")


@pytest.mark.gpu
def test_backend_film_matches_pass_abi(device):
    s = cornell()
    rep, film = s.backend_render(48, 32, 4, max_depth=6)
    assert rep["ok"] == 1 and rep["error"] == "", rep["error"]
    assert rep["progress_calls"] == 4 and rep["last_progress_spp"] == 4
    assert rep["samples_rendered"] == 48 * 32 * 4
    s.build_accel()
    ref, st = device.render(s, 48, 32, 4, max_depth=6)
    ref = ref.reshape(32, 48)
    assert np.array_equal(film["samples"], ref["samples"])
    assert np.array_equal(film["sum"], ref["sum"].astype(np.float32))
    assert np.array_equal(film["aov_samples"], ref["aov_samples"])
    assert rep["rays_traced"] == st["rays_traced"] and rep["hits"] == st["hits"]
    assert rep["bvh_nodes"] == s.info.node_count and rep["threads"] > 0


@pytest.mark.gpu
def test_progress_cancel_and_resume(device):
    s = cornell()
    full, ffull = s.backend_render(32, 24, 5, max_depth=5)
    assert full["ok"] == 1
    part, fpart = s.backend_render(32, 24, 5, max_depth=5, cancel_after_passes=2)
    assert part["ok"] == 0 and part["error"] == "render cancelled by progress callback"
    assert fpart is not None and np.all(fpart["samples"] == 2)
    named, _ = s.backend_render(32, 24, 5, max_depth=5, cancel_after_passes=1, cancel_error="stop here")
    assert named["error"] == "stop here"
    done, fdone = s.backend_render(32, 24, 5, max_depth=5, resume_film=fpart, resume_completed_spp=2)
    assert done["ok"] == 1 and done["progress_calls"] == 3
    assert np.all(fdone["samples"] == 5) and done["samples_rendered"] == 32 * 24 * 5
    # The resumed film restarts from the FP32 checkpoint, so sums agree to
    # FP32 rounding of the first two passes.
    assert np.allclose(fdone["sum"], ffull["sum"], rtol=1e-5, atol=1e-5)
    assert np.array_equal(fdone["aov_samples"], ffull["aov_samples"])


@pytest.mark.gpu
def test_resume_validation_and_stop(device):
    s = cornell()
    rep, film = s.backend_render(16, 16, 4, stop_before_start=True)
    assert rep["ok"] == 0 and rep["error"] == "render cancelled"
    assert film is not None and np.all(film["samples"] == 0)
    rep, _ = s.backend_render(16, 16, 4, resume_completed_spp=5)
    assert rep["error"] == "invalid resume completed spp"
    bad = np.zeros((16, 8), FILM32_DTYPE)
    rep, _ = s.backend_render(16, 16, 4, resume_film=bad, resume_completed_spp=0)
    assert rep["error"] == "resume film dimensions do not match render scene"
    wrong = np.zeros((16, 16), FILM32_DTYPE)
    wrong["samples"] = 1
    rep, _ = s.backend_render(16, 16, 4, resume_film=wrong, resume_completed_spp=2)
    assert rep["error"] == "resume film sample counts do not match resume completed spp"
    rep, film = s.backend_render(16, 16, 4, resume_film=wrong, resume_completed_spp=1)
    assert rep["ok"] == 1 and np.all(film["samples"] == 4)
    assert rep["progress_calls"] == 3
