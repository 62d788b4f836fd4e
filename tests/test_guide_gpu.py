"""GPU parity of the learned guiding proposal (BF16 MFMA kernels) against the
FP64 oracle, and unbiasedness of guided rendering.

* network evaluation and gradients are compared with the oracle run in its
  BF16-emulating mode (same rounding points as the device), so the remaining
  differences are FP32 accumulation order: tolerances are stated per check;
* the device gradients also agree in direction with the plain double
  gradients (which finite differences pin in test_oracle_guide.py);
* mixture pdf and sampling equal the oracle's within FP32 rounding;
* training on samples drawn from a known lobe concentrates the proposal on it;
* a guided render keeps the closed-form answer (diffuse 0.8 under a constant
  environment) and the plain integrator's mean on the Cornell box.
"""
import numpy as np
import pytest

from oracle import oracle_py as O
from tests.helpers import base_path_scene, guide_desc, guide_params, guide_samples
from yaoray_amd import Scene
from yaoray_amd import native as N

pytestmark = pytest.mark.gpu


def configure(device, d):
    device.guide_configure(list(d.bounds_min), list(d.bounds_max), d.log2_table_size, d.min_resolution,
                           d.max_resolution, d.learning_rate, d.guide_fraction, d.sample_capacity, d.seed)


def test_guide_config_matches_oracle(device):
    for log2, rmin, rmax in ((12, 4, 64), (16, 8, 256)):
        d = guide_desc(log2, rmin, rmax)
        configure(device, d)
        m, g, res, dense = device.guide_info()
        ores, odense = O.guide_config(d)
        assert m == N.GUIDE_MLP_PARAMS and g == 8 * (1 << log2) * 2
        assert np.array_equal(res, ores) and np.array_equal(dense, odense)


def test_query_matches_bf16_oracle(device):
    d = guide_desc()
    configure(device, d)
    mlp, grid = guide_params(d, 3)
    device.guide_set_params(mlp, grid)
    rows = guide_samples(1000, 21)[:, :9]
    g = device.guide_query(rows)
    o = O.guide_forward(d, mlp, grid, rows, bf16=True)
    dp = np.abs(g[:, :8] - o[:, :8])
    dk = np.abs(g[:, 8:16] - o[:, 8:16]) / o[:, 8:16]
    dm = np.abs(g[:, 16:] - o[:, 16:])
    print(f"query: pi {dp.max():.2e} kappa {dk.max():.2e} mu {dm.max():.2e}")
    assert np.percentile(dp, 99) < 2e-3 and dp.max() < 2e-2
    assert np.percentile(dk, 99) < 5e-3 and dk.max() < 5e-2
    assert np.percentile(dm, 99) < 5e-3 and dm.max() < 5e-2
    # Against plain double: BF16 rounding only.
    od = O.guide_forward(d, mlp, grid, rows)
    assert np.abs(g[:, :8] - od[:, :8]).max() < 0.1


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def test_gradients_match_bf16_oracle(device):
    d = guide_desc()
    configure(device, d)
    mlp, grid = guide_params(d, 3)
    device.guide_set_params(mlp, grid)
    smp = guide_samples(777, 31)  # not a multiple of 16: a partial row block
    gm, gg, loss = device.guide_gradients(smp, 0.8)
    om, og, oloss = O.guide_gradients(d, mlp, grid, smp, 0.8, bf16=True)
    bounds = [0, 2048, 6144, 10240, 13312]
    errs = [rel_l2(gm[a:b], om[a:b]) for a, b in zip(bounds[:-1], bounds[1:])]
    gerr = rel_l2(gg, og)
    print(f"gradients: layers {['%.2e' % e for e in errs]} grid {gerr:.2e} loss {loss:.6f} / {oloss:.6f}")
    assert max(errs) < 2e-2 and gerr < 2e-2
    assert abs(loss - oloss) < 1e-3 * abs(oloss)
    # Direction agrees with the double gradient (finite-difference pinned).
    dm, dg, _ = O.guide_gradients(d, mlp, grid, smp, 0.8)
    cos = float(gm @ dm / (np.linalg.norm(gm) * np.linalg.norm(dm)))
    assert cos > 0.99, cos
    assert np.all(gm[10240:].reshape(64, 48)[:, 40:] == 0.0)


def test_mixture_eval_matches_oracle(device):
    rng = np.random.default_rng(41)
    n = 20000
    mix = np.zeros((n, 40), np.float32)
    mix[:, :8] = rng.dirichlet(np.ones(8), n)
    mix[:, 8:16] = np.exp(rng.uniform(-3, 6.2, (n, 8)))
    mu = rng.normal(size=(n, 8, 3))
    mu /= np.linalg.norm(mu, axis=2, keepdims=True)
    mix[:, 16:] = mu.reshape(n, 24)
    dirs = rng.normal(size=(n, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    u = rng.uniform(size=(n, 3)).astype(np.float32)
    gp, gs = device.guide_mixture_eval(mix, dirs, u)
    op, os_ = O.guide_mixture_eval(mix, dirs, u)
    assert np.allclose(gp, op, rtol=2e-3, atol=1e-6)
    ang = np.abs(gs[:, :3] - os_[:, :3]).max(axis=1)
    assert np.percentile(ang, 99.9) < 1e-3
    assert np.allclose(np.linalg.norm(gs[:, :3], axis=1), 1.0, atol=1e-5)


def test_training_learns_a_lobe(device):
    # A coarse grid: the target does not depend on position, and fine hashed
    # levels would memorise the training positions instead of generalising.
    d = guide_desc(log2_table=12, rmin=2, rmax=8, lr=3e-3)
    configure(device, d)
    rng = np.random.default_rng(51)
    n = 60000
    target = np.array([0.0, 1.0, 0.0])
    # Directions from a vMF(kappa=30) lobe around +y: inversion in the lobe frame.
    kappa = 30.0
    u1, u2 = rng.uniform(size=n), rng.uniform(size=n)
    w = 1.0 + np.log(u1 + (1 - u1) * np.exp(-2 * kappa)) / kappa
    st = np.sqrt(np.maximum(0, 1 - w * w))
    wi = np.stack([st * np.cos(2 * np.pi * u2), w, st * np.sin(2 * np.pi * u2)], axis=1)
    smp = guide_samples(n, 52)
    smp[:, 9:12] = wi
    smp[:, 12] = 1.0
    probe = smp[rng.choice(n, 2000, replace=False), :9]
    before = device.guide_query(probe)
    p0, _ = device.guide_mixture_eval(before, np.tile(target, (len(probe), 1)), np.zeros((len(probe), 3)))
    _, _, loss0 = device.guide_gradients(smp[:4096], 1.0)
    device.guide_add_samples(smp)
    count, mean_w = device.guide_sample_stats()
    assert count == n and abs(mean_w - 1.0) < 1e-5
    device.guide_train(400, 4096)
    assert device.guide_sample_stats()[0] == 0
    after = device.guide_query(probe)
    p1, _ = device.guide_mixture_eval(after, np.tile(target, (len(probe), 1)), np.zeros((len(probe), 3)))
    _, _, loss1 = device.guide_gradients(smp[:4096], 1.0)
    true_pdf = kappa / (2 * np.pi * (1 - np.exp(-2 * kappa)))
    print(f"training: pdf at lobe {np.median(p0):.3f} -> {np.median(p1):.3f} (true {true_pdf:.2f}), "
          f"loss {loss0:.4f} -> {loss1:.4f}")
    assert np.median(p1) > 0.5 * true_pdf and np.median(p1) > 10 * np.median(p0)
    assert loss1 < loss0 - 1.0


def guided_closed_form(device, mode):
    s = base_path_scene()
    device.upload_scene(s)
    device.upload_shading(s)
    configure(device, guide_desc(bounds=((-2, -2, -1), (2, 2, 1))))
    device.guide_set_mode(mode)
    try:
        film, st = device.render(s, 9, 9, 256, max_depth=2, sampler=2, seed=3, upload=False)
    finally:
        device.guide_set_mode(N.GUIDE_OFF)
    return film, st


@pytest.mark.parametrize("mode", [N.GUIDE_SAMPLE, N.GUIDE_RECORD])
def test_guided_render_keeps_closed_form(device, mode):
    film, _ = guided_closed_form(device, mode)
    env = np.array([0.02, 0.03, 0.04])
    f = film.reshape(9, 9)
    centre = f[3:6, 3:6].reshape(-1)
    mean = centre["sum"].sum(axis=0) / centre["samples"].sum()
    lum_var = (centre["luminance_m2"].sum()) / max(centre["samples"].sum() - 1, 1)
    stderr = np.sqrt(lum_var / centre["samples"].sum())
    lum = lambda c: float(np.dot(c, [0.2126, 0.7152, 0.0722]))
    print(f"guided closed form: {mean} vs {0.8 * env}, stderr {stderr:.2e}")
    assert stderr > 0.0  # the guided estimator is not the cosine-exact one
    assert abs(lum(mean) - lum(0.8 * env)) < 4 * stderr
    assert np.allclose(mean, 0.8 * env, rtol=0.02)
    if mode == N.GUIDE_RECORD:
        count, mean_w = device.guide_sample_stats()
        assert count > 0 and mean_w > 0.0


@pytest.mark.parametrize("depth_limit", [0, 2])
def test_guided_cornell_matches_plain_mean(device, depth_limit):
    """Guided and plain renders have the same expectation, with every
    continuing vertex guided or only those of the first two bounces (plain
    vertices after guided ones). Per-pixel sample
    variances understate the error of scrambled-Sobol pixel means, so the
    error is estimated from paired differences over 16 independent sampler
    seeds (a t statistic with 15 degrees of freedom, two-sided 0.1 % bound:
    the paired differences are heavy-tailed, and fewer seeds made the
    bound trip on unbiased runs). A 1 % bias gives |t| above 10 here."""
    s = Scene.cornell(1)
    s.build_accel()
    w, h, spp = 64, 64, 16
    configure(device, guide_desc(log2_table=16, rmin=8, rmax=128, bounds=((-1.05, -0.05, -1.05), (1.05, 2.05, 1.05))))
    device.upload_scene(s)
    device.upload_shading(s)
    device.guide_set_depth_limit(depth_limit)
    device.guide_set_mode(N.GUIDE_RECORD)
    try:
        for k in range(4):  # a few training rounds on the integrator's own samples
            device.render(s, w, h, 4, max_depth=8, upload=False, seed=900 + k)
            device.guide_train(32, 8192)
    finally:
        device.guide_set_mode(N.GUIDE_OFF)
    lum = np.array([0.2126, 0.7152, 0.0722])
    diffs, means = [], []
    for seed in range(16):
        plain, _ = device.render(s, w, h, spp, max_depth=8, upload=False, seed=seed)
        device.guide_set_mode(N.GUIDE_SAMPLE)
        try:
            guided, _ = device.render(s, w, h, spp, max_depth=8, upload=False, seed=seed)
        finally:
            device.guide_set_mode(N.GUIDE_OFF)
        gl = ((guided["sum"] / guided["samples"][:, None]) @ lum).mean()
        pl = ((plain["sum"] / plain["samples"][:, None]) @ lum).mean()
        diffs.append(gl - pl)
        means.append(pl)
    diffs = np.array(diffs)
    t = diffs.mean() / (diffs.std(ddof=1) / np.sqrt(len(diffs)))
    print(f"cornell guided - plain: {diffs.mean():.5f} +- {diffs.std(ddof=1) / np.sqrt(len(diffs)):.5f} "
          f"(plain {np.mean(means):.5f}), t = {t:.2f}")
    assert abs(t) < 4.073  # two-sided 0.1 % for 15 degrees of freedom
    assert abs(diffs.mean()) < 0.01 * np.mean(means)


def test_guided_room_converges_to_64k_spp_reference(device):
    """Against a 65536-spp FP64 plain reference of the window-lit room, the
    guided render at 4096 spp is closer (relative MSE) than the plain one at
    the same spp, and both image means are near the reference's. The room's
    sun-lit glossy paths make the per-pixel noise heavy-tailed (image means
    move by ~2 % between sampler seeds at 4096 spp), so the mean bound is 5 %;
    unbiasedness itself is tested on the Cornell box and the closed form."""
    s = Scene.room()
    s.build_accel()
    device.upload_scene(s)
    device.upload_shading(s)
    w, h, depth = 48, 27, 8
    lum = np.array([0.2126, 0.7152, 0.0722])
    ref, _ = device.render(s, w, h, 65536, max_depth=depth, upload=False, batch=4096, seed=77)
    device.guide_configure([-0.1, -0.1, -0.1], [10.1, 4.1, 10.1], 18, 8, 512, 3e-3, 0.5, 1 << 20, 7)
    device.film_reset(w, h)
    films = {}
    try:
        device.guide_set_mode(N.GUIDE_RECORD)
        for k in range(8):
            device.render_pass(s.render_desc(w, h, 4096, depth, seed=79, sample_index=k))
            device.guide_train(16, 8192)
        for mode in (N.GUIDE_OFF, N.GUIDE_SAMPLE):
            device.guide_set_mode(mode)
            device.film_reset(w, h)
            d = s.render_desc(w, h, 4096, depth, seed=100)
            d.sample_count = 4096
            device.render_pass(d)
            device.synchronize()
            films[mode] = device.film_read(w, h)
    finally:
        device.guide_set_mode(N.GUIDE_OFF)
    img = lambda f: f["sum"] / np.maximum(f["samples"], 1)[:, None]
    r = img(ref)
    rel = lambda x: float(np.mean((x - r) ** 2 / (r ** 2 + 0.01)))
    g, p = img(films[N.GUIDE_SAMPLE]), img(films[N.GUIDE_OFF])
    rl = (r @ lum).mean()
    print(f"room @4096 spp: relMSE guided {rel(g):.5f} plain {rel(p):.5f}; mean guided {(g @ lum).mean():.5f} "
          f"plain {(p @ lum).mean():.5f} reference {rl:.5f}")
    assert rel(g) < 0.9 * rel(p)
    assert abs((g @ lum).mean() / rl - 1.0) < 0.05
    assert abs((p @ lum).mean() / rl - 1.0) < 0.05


def test_cell_copies_leave_guided_render_unchanged(device):
    """Per-cell copies of the grid change how sampling passes read the
    features, not their values: a guided render of the room is bit-identical
    with no copies, the default budget (dense and coarse hashed levels
    copied) and a budget that copies every level."""
    s = Scene.room()
    s.build_accel()
    w, h = 160, 90
    device.guide_configure([-0.1, -0.1, -0.1], [10.1, 4.1, 10.1], log2_table_size=18, min_resolution=8,
                           max_resolution=512, learning_rate=3e-3, guide_fraction=0.5, sample_capacity=1 << 18,
                           seed=7)
    device.upload_scene(s)
    device.upload_shading(s)
    device.guide_set_mode(N.GUIDE_RECORD)
    try:
        for k in range(2):
            device.render(s, w, h, 4, max_depth=8, upload=False, seed=700 + k)
            device.guide_train(16, 8192)
        films = []
        for budget in (0, 512 << 20, 16 << 30, 0):
            device.guide_set_cell_budget(budget)
            device.guide_set_mode(N.GUIDE_SAMPLE)
            film, _ = device.render(s, w, h, 8, max_depth=8, upload=False, seed=11, batch=8)
            films.append(film)
    finally:
        device.guide_set_mode(N.GUIDE_OFF)
        device.guide_set_cell_budget(512 << 20)
    assert films[0]["samples"].sum() == w * h * 8
    for f in films[1:]:
        assert np.array_equal(f.view(np.uint8), films[0].view(np.uint8))
