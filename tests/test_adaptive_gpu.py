"""Adaptive sampling on the device (yrh_film_update_active + pixel-list passes)
against the reference's adaptive RenderCpuPathTrace (golden fixture
cornell_adaptive, written by oracle/_ref) and against closed cases.

The device keeps its film in FP64 and applies Film::IsConverged to the
statistics rounded to FP32, while the reference runs the test on its FP32
Welford sums: a pixel whose error bound lies within rounding of the
tolerance can stop one pass earlier or later, and depth-8 paths drift where
the device's transcendental functions differ from glibc in the last bit. So
the per-pixel sample counts are compared as a fraction of equal pixels
(all at depth 1, whose paths use only + - * / sqrt; at least 99 % at depth
8 — on the fixture all are equal) and the converged counts within a stated
tolerance; pixels with equal counts are compared value by value.
"""
import numpy as np
import pytest

from tests import golden_data as G
from yaoray_amd import Scene

pytestmark = pytest.mark.gpu


def adaptive_args(f, depth):
    a = [float(v) for v in f[f"adaptive{depth}"]]
    return (int(a[0]), a[1], a[2], a[3])


@pytest.mark.parametrize("depth,min_equal", [(1, 1.0), (8, 1.0)])
def test_adaptive_matches_reference_fixture(device, depth, min_equal):
    f = G.load("cornell_adaptive")
    s = G.cornell_scene(f)
    w, h = int(f["width"]), int(f["height"])
    spp = int(f[f"spp{depth}"])
    film, st, passes, conv = device.render_adaptive(s, w, h, spp, adaptive_args(f, depth), max_depth=depth,
                                                    sampler=2, seed=int(f["seed"]))
    ref = G.film(f[f"film{depth}"])
    equal = film["samples"] == ref["samples"]
    print(f"depth {depth}: equal counts {equal.mean():.4f}, converged {conv} vs {int(f[f'converged{depth}'])}, "
          f"samples {film['samples'].sum()} vs {ref['samples'].sum()}, passes {passes}")
    assert equal.mean() >= min_equal
    assert abs(conv - int(f[f"converged{depth}"])) <= 0.02 * w * h
    assert abs(int(film["samples"].sum()) - int(ref["samples"].sum())) <= 0.01 * int(ref["samples"].sum())
    assert passes == spp
    # Where the counts agree the pixel means agree (bit-equal sums rounded to
    # FP32 at depth 1; at depth 8 within the drift of the plain render).
    n = film["samples"][equal].astype(np.float64)[:, None]
    mean_dev = film["sum"][equal] / n
    mean_ref = ref["sum"][equal].astype(np.float64) / n
    tol = 4 * 2.0 ** -24 if depth == 1 else 1e-3
    assert np.allclose(mean_dev, mean_ref, rtol=tol, atol=tol)


def test_adaptive_closed_cases(device):
    """Adaptive sampling off is the plain render, bit for bit; a tolerance no
    pixel can miss stops every pixel at max(2, min_spp); a zero tolerance
    stops exactly the pixels whose luminance never varied, at 2 samples."""
    s = Scene.cornell(1)
    s.build_accel()
    plain, pst = device.render(s, 24, 24, 6, max_depth=4, batch=6)
    off, ost, passes, conv = device.render_adaptive(s, 24, 24, 6, None, max_depth=4, upload=False)
    assert off.tobytes() == plain.tobytes() and passes == 6 and conv == 0
    assert ost["rays_traced"] == pst["rays_traced"]
    loose, _, passes, conv = device.render_adaptive(s, 24, 24, 12, (3, 1e30, 1e30, 1.96), max_depth=4, upload=False)
    assert np.all(loose["samples"] == 3) and passes == 3 and conv == 24 * 24
    tight, _, _, conv = device.render_adaptive(s, 24, 24, 8, (2, 0.0, 0.0, 1.96), max_depth=4, upload=False)
    still = tight["luminance_m2"] == 0.0
    assert np.all(tight["samples"][still] == 2) and np.all(tight["samples"][~still] == 8)
    assert conv == int(still.sum())


def test_pixel_list_pass_needs_a_list(device):
    s = Scene.cornell(1)
    s.build_accel()
    device.upload_scene(s)
    device.upload_shading(s)
    device.film_reset(8, 8)
    d = s.render_desc(8, 8, 4, 4, 2, 1, 0.0, 0)
    d.pixel_list = 1
    with pytest.raises(Exception, match="active list"):
        device.render_pass(d)


def test_backend_adaptive_render_and_resume(device):
    """yr::HipBackend with adaptive sampling: the same per-pixel sample counts
    as the device loop, the converged count in the stats and the progress
    reports, and the reference's resume rule (converged pixels may hold fewer
    samples than the completed passes only when adaptive sampling is on)."""
    f = G.load("cornell_adaptive")
    s = G.cornell_scene(f)
    a = adaptive_args(f, 8)
    film, _, passes, conv = device.render_adaptive(s, 64, 64, 32, a, max_depth=8, sampler=2, seed=int(f["seed"]))
    rep, bfilm = s.backend_render(64, 64, 32, max_depth=8, sampler=2, seed=int(f["seed"]), adaptive=a)
    assert rep["ok"] == 1, rep["error"]
    bs = bfilm["samples"].reshape(-1)
    ref = G.film(f["film8"])
    print(f"backend: samples {int(bs.sum())} device loop {int(film['samples'].sum())} reference {int(ref['samples'].sum())}; "
          f"differing pixels {int((bs != film['samples']).sum())}; converged {rep['converged_pixels']} vs {conv}; "
          f"passes {passes}, progress calls {rep['progress_calls']}, rays {rep['rays_traced']}")
    assert np.array_equal(bs, film["samples"])
    assert rep["converged_pixels"] == conv and rep["samples_rendered"] == int(film["samples"].sum())
    # Interrupted after 8 passes (by the progress callback) and resumed.
    part, fpart = s.backend_render(64, 64, 32, max_depth=8, sampler=2, seed=int(f["seed"]), adaptive=a,
                                   cancel_after_passes=8)
    assert part["ok"] == 0 and part["last_progress_spp"] == 8
    assert fpart["samples"].max() == 8 and fpart["samples"].min() >= a[0]
    plain_resume, _ = s.backend_render(64, 64, 32, max_depth=8, sampler=2, seed=int(f["seed"]), resume_film=fpart,
                                       resume_completed_spp=8)
    if fpart["samples"].min() < 8:
        assert plain_resume["error"] == "resume film sample counts do not match resume completed spp"
    done, fdone = s.backend_render(64, 64, 32, max_depth=8, sampler=2, seed=int(f["seed"]), adaptive=a,
                                   resume_film=fpart, resume_completed_spp=8)
    assert done["ok"] == 1, done["error"]
    same = fdone["samples"].reshape(-1) == film["samples"]
    print(f"resumed adaptive: equal counts {same.mean():.4f}, last active {done['last_progress_active']}")
    assert same.mean() == 1.0  # resumed from the FP32 film: the same counts on every pixel


@pytest.mark.parametrize("batch", [4, 16])
def test_batched_adaptive_passes_equal_one_sample_passes(device, batch):
    """A pixel-list pass of several samples per pixel, accepted in order until
    the pixel converges, leaves the same film (bytes) and sample counts as the
    one-sample passes; its counters include the samples it did not add."""
    f = G.load("cornell_adaptive")
    s = G.cornell_scene(f)
    a = adaptive_args(f, 8)
    one, st1, passes1, conv1 = device.render_adaptive(s, 64, 64, 32, a, max_depth=8, sampler=2, seed=int(f["seed"]))
    many, stk, passesk, convk = device.render_adaptive(s, 64, 64, 32, a, max_depth=8, sampler=2, seed=int(f["seed"]),
                                                       upload=False, batch=batch)
    assert many.tobytes() == one.tobytes()
    assert convk == conv1 == int(f["converged8"]) and passesk == passes1
    assert np.array_equal(many["samples"], G.film(f["film8"])["samples"])
    assert stk["rays_traced"] >= st1["rays_traced"] and stk["bad_samples"] == st1["bad_samples"]


def test_backend_without_progress_callback(device):
    """Without a progress callback HipBackend batches passes: plain renders in
    one wavefront of many sample indices, adaptive ones in 16-sample pixel-list
    passes. Both leave the films of the one-pass-per-sample renders."""
    f = G.load("cornell_adaptive")
    s = G.cornell_scene(f)
    a = adaptive_args(f, 8)
    seed = int(f["seed"])
    rep1, film1 = s.backend_render(64, 64, 32, max_depth=8, sampler=2, seed=seed, adaptive=a)
    repb, filmb = s.backend_render(64, 64, 32, max_depth=8, sampler=2, seed=seed, adaptive=a, cancel_after_passes=-1)
    assert rep1["ok"] == 1 and repb["ok"] == 1 and repb["progress_calls"] == 0
    assert filmb.tobytes() == film1.tobytes()
    assert repb["converged_pixels"] == rep1["converged_pixels"] == int(f["converged8"])
    assert repb["rays_traced"] >= rep1["rays_traced"]
    p1, pf1 = s.backend_render(48, 32, 6, max_depth=6, seed=seed)
    pb, pfb = s.backend_render(48, 32, 6, max_depth=6, seed=seed, cancel_after_passes=-1)
    assert pb["ok"] == 1 and pb["progress_calls"] == 0
    assert pfb.tobytes() == pf1.tobytes() and pb["rays_traced"] == p1["rays_traced"]


def test_backend_adaptive_passes_respect_the_path_cap(device, monkeypatch):
    """Adaptive passes without a progress callback are capped at the paths a pass
    may hold (device memory; here forced down with YAORAY_HIP_MAX_PATHS_PER_PASS):
    min_spp x pixels is 16x the cap, so the min_spp samples are split over many
    pixel-list passes. The film is the uncapped one, byte for byte."""
    f = G.load("cornell_adaptive")
    s = G.cornell_scene(f)
    a = (16, 0.1, 0.001, 1.96)
    seed = int(f["seed"])
    rep, film = s.backend_render(64, 64, 24, max_depth=8, sampler=2, seed=seed, adaptive=a, cancel_after_passes=-1)
    monkeypatch.setenv("YAORAY_HIP_MAX_PATHS_PER_PASS", str(64 * 64))
    repc, filmc = s.backend_render(64, 64, 24, max_depth=8, sampler=2, seed=seed, adaptive=a, cancel_after_passes=-1)
    assert rep["ok"] == 1 and repc["ok"] == 1, (rep["error"], repc["error"])
    assert filmc.tobytes() == film.tobytes()
    assert repc["converged_pixels"] == rep["converged_pixels"]
    assert repc["rays_traced"] <= rep["rays_traced"]  # one-sample passes discard no samples
