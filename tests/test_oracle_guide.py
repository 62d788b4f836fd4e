"""The FP64 oracle of the learned guiding proposal, pinned without the device:

* its gradients equal central finite differences of its own loss (a dropped
  term, a wrong sign or a transposed weight index fails this);
* its grid configuration matches the closed form of the level resolutions;
* every vMF lobe integrates to one over the sphere (quadrature in the lobe's
  own polar angle), and a mixture integrates to one;
* sampling follows the pdf: a histogram of sampled directions matches the
  pdf integrated over the same bins.
"""
import numpy as np
import pytest

from oracle import oracle_py as O
from tests.helpers import guide_desc as desc
from tests.helpers import guide_params as random_params
from tests.helpers import guide_samples as random_samples


def test_grid_configuration_closed_form():
    for log2, rmin, rmax in ((12, 4, 64), (16, 8, 256), (19, 16, 1024)):
        d = desc(log2, rmin, rmax)
        res, dense = O.guide_config(d)
        for lvl in range(8):
            expect = int(np.floor(rmin * (rmax / rmin) ** (lvl / 7.0) + 0.5))
            assert res[lvl] == expect
            assert dense[lvl] == int((expect + 1) ** 3 <= (1 << log2))
        assert res[0] == rmin and res[7] == rmax


def test_gradients_match_finite_differences():
    d = desc()
    mlp, grid = random_params(d, 3)
    smp = random_samples(6, 4)
    gm, gg, loss = O.guide_gradients(d, mlp, grid, smp, 0.9)
    assert np.isfinite(loss) and loss != 0.0
    assert abs(O.guide_loss(d, mlp, grid, smp, 0.9) - loss) < 1e-12 * max(1.0, abs(loss))
    rng = np.random.default_rng(5)
    # MLP weights with a non-negligible gradient in each layer, plus random ones.
    bounds = [0, 2048, 6144, 10240, 13312]
    picks = []
    for a, b in zip(bounds[:-1], bounds[1:]):
        seg = np.arange(a, b)
        seg = seg[np.abs(gm[a:b]) > 1e-6 * np.abs(gm).max()]
        picks.extend(rng.choice(seg, size=min(12, len(seg)), replace=False))
    touched = np.flatnonzero(gg)
    assert len(touched) > 0
    grid_picks = rng.choice(touched, size=min(24, len(touched)), replace=False)
    h = 1.0e-4
    bad = 0
    checked = 0
    for which, idxs, base, grad in (("mlp", picks, mlp, gm), ("grid", grid_picks, grid, gg)):
        for i in idxs:
            plus, minus = base.copy(), base.copy()
            plus[i] += np.float32(h)
            minus[i] -= np.float32(h)
            step = float(plus[i]) - float(minus[i])
            args_p = (plus, grid) if which == "mlp" else (mlp, plus)
            args_m = (minus, grid) if which == "mlp" else (mlp, minus)
            fd = (O.guide_loss(d, *args_p, smp, 0.9) - O.guide_loss(d, *args_m, smp, 0.9)) / step
            checked += 1
            if abs(fd - grad[i]) > 1e-4 * max(abs(fd), abs(grad[i])) + 1e-9:
                bad += 1
    # A ReLU kink inside [x - h, x + h] can spoil an isolated difference.
    assert checked > 40 and bad <= 2, (bad, checked)


def test_gradient_of_a_dropped_term_is_detected():
    """The check above has teeth: zeroing the weight of one sample changes the
    gradient by far more than the tolerance."""
    d = desc()
    mlp, grid = random_params(d, 3)
    smp = random_samples(6, 4)
    gm, _, _ = O.guide_gradients(d, mlp, grid, smp, 0.9)
    smp2 = smp.copy()
    smp2[0, 12] = 0.0
    gm2, _, _ = O.guide_gradients(d, mlp, grid, smp2, 0.9)
    assert np.abs(gm - gm2).max() > 1e-2 * np.abs(gm).max()


def one_lobe(kappa, mu=(0.0, 0.0, 1.0)):
    m = np.zeros(40, np.float32)
    m[0] = 1.0
    m[8:16] = 1.0
    m[8] = kappa
    for k in range(8):
        m[16 + 3 * k:16 + 3 * k + 3] = (0.0, 0.0, 1.0)
    m[16:19] = mu
    return m


@pytest.mark.parametrize("kappa", [0.05, 1.0, 10.0, 100.0, 490.0])
def test_vmf_lobe_integrates_to_one(kappa):
    theta = np.linspace(0.0, np.pi, 400001)
    dirs = np.stack([np.sin(theta), np.zeros_like(theta), np.cos(theta)], axis=1)
    mix = np.tile(one_lobe(kappa), (len(theta), 1))
    pdf, _ = O.guide_mixture_eval(mix, dirs, np.zeros((len(theta), 3)))
    integral = np.trapezoid(pdf.astype(np.float64) * 2 * np.pi * np.sin(theta), theta)
    assert abs(integral - 1.0) < 2e-4


def random_mixture(seed):
    rng = np.random.default_rng(seed)
    m = np.zeros(40, np.float32)
    pi = rng.dirichlet(np.ones(8))
    m[:8] = pi
    m[8:16] = np.exp(rng.uniform(-2, 4, 8))
    mu = rng.normal(size=(8, 3))
    mu /= np.linalg.norm(mu, axis=1, keepdims=True)
    m[16:] = mu.ravel()
    return m


def test_mixture_integrates_to_one_and_sampling_follows_pdf():
    m = random_mixture(11)
    n = 400000
    rng = np.random.default_rng(12)
    # Integral by uniform sphere sampling.
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    pdf, _ = O.guide_mixture_eval(np.tile(m, (n, 1)), v, np.zeros((n, 3)))
    est = pdf.astype(np.float64).mean() * 4 * np.pi
    err = pdf.astype(np.float64).std() * 4 * np.pi / np.sqrt(n)
    assert abs(est - 1.0) < 5 * err + 1e-3
    # Samples: histogram over octant x 4 polar bands vs the pdf mass per bin.
    u = rng.uniform(size=(n, 3)).astype(np.float32)
    _, smp = O.guide_mixture_eval(np.tile(m, (n, 1)), np.zeros((n, 3)), u)
    dirs = smp[:, :3].astype(np.float64)
    assert np.allclose(np.linalg.norm(dirs, axis=1), 1.0, atol=1e-5)
    pdf_at, _ = O.guide_mixture_eval(np.tile(m, (n, 1)), dirs, np.zeros((n, 3)))
    assert np.allclose(pdf_at, smp[:, 3], rtol=1e-5, atol=1e-7)

    def bins(d):
        octant = (d[:, 0] > 0) * 1 + (d[:, 1] > 0) * 2 + (d[:, 2] > 0) * 4
        band = np.clip(((np.abs(d[:, 2])) * 4).astype(int), 0, 3)
        return octant * 4 + band

    observed = np.bincount(bins(dirs), minlength=32) / n
    expected = np.bincount(bins(v), weights=pdf.astype(np.float64) * 4 * np.pi, minlength=32) / n
    sigma = np.sqrt(observed * (1 - observed) / n) + np.sqrt(expected / n) * 0.5 + 1e-4
    assert np.all(np.abs(observed - expected) < 5 * sigma + 2e-3), np.abs(observed - expected).max()


def test_bf16_emulation_is_close_to_double():
    """The BF16 evaluation differs from the double one by BF16 rounding only."""
    d = desc()
    mlp, grid = random_params(d, 3)
    rows = random_samples(64, 9)[:, :9]
    a = O.guide_forward(d, mlp, grid, rows)
    b = O.guide_forward(d, mlp, grid, rows, bf16=True)
    assert np.allclose(a[:, :8], b[:, :8], atol=5e-2)
    assert not np.array_equal(a, b)
