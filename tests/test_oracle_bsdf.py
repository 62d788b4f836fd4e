"""CPU pins of the oracle's BSDFs (EvaluateBsdf / PdfBsdf / SampleBsdf /
IsDeltaBsdf, src/shading/bsdf.cpp:17-329, bsdf_microfacet.cpp:20-298,
bsdf_common.cpp:13-47), branch by branch:

* the reference's own known answers (tests/shading/bsdf_diffuse_conductor_tests.cpp:12-116,
  bsdf_dielectric_tests.cpp:10-146);
* bit equality with the reference build (oracle/_ref) on random queries over
  every branch, and with the golden fixture it wrote (tests/golden/bsdf.npz);
* what the mathematics fixes where it holds for the reference's formulation:
  Lambert's closed forms; for GGX conductors, the pdf integrates to the
  probability of a valid sample and the sampling estimator's mean equals the
  quadrature of f cos (a wrong Jacobian, D, G or Fresnel term breaks one);
  reflection reciprocity of conductors and rough dielectrics; the sampled pdf
  and weight equal PdfBsdf and f |cos| / pdf at the sampled direction; smooth
  dielectrics reflect with the Fresnel probability and carry weight
  reflectance. (The rough dielectric's sampler picks reflection with the
  Fresnel term of a different half vector than the one it keeps, so its mean
  weight is not the quadrature of f cos in the reference either; those
  branches are pinned by the reference build and the fixture.)
"""
import numpy as np
import pytest

from oracle import oracle_py as O
from tests import bsdf_cases as B
from tests import golden_data as G

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (reference sources absent)")

Z = np.array([0.0, 0.0, 1.0])


def run_known_answers(query):
    res = {}
    for name, m, row in B.known_answer_cases():
        res[name] = query([m], [0], row[None])[0]
    return res


def test_oracle_bsdf_known_answers():
    B.check_known_answers(run_known_answers(O.bsdf_query))


@needs_ref
def test_reference_bsdf_known_answers():
    # The cases restate the reference's tests; the reference build passes them too.
    B.check_known_answers(run_known_answers(O.ref_bsdf_query))


@needs_ref
def test_oracle_bsdf_matches_reference_bits():
    mi, q = B.queries(2000)
    a = O.bsdf_query(B.MATERIALS, mi, q)
    b = O.ref_bsdf_query(B.MATERIALS, mi, q)
    assert a.view(np.uint32).tobytes() == b.view(np.uint32).tobytes()
    flags = a[:, 11].astype(int)
    # every branch is exercised: valid glossy, valid specular, invalid samples
    for m in range(len(B.MATERIALS)):
        f = flags[mi == m]
        assert (f & 1).any(), m


def test_oracle_bsdf_matches_fixture():
    f = G.load("bsdf")
    mats = G.bsdf_materials(f)
    out = O.bsdf_query(mats, f["material_index"], f["queries"])
    assert out.view(np.uint32).tobytes() == f["results"].view(np.uint32).tobytes()


def sphere_grid(nt=300, nphi=600):
    ct = -1.0 + (np.arange(2 * nt) + 0.5) / nt
    ph = (np.arange(nphi) + 0.5) / nphi * 2.0 * np.pi
    c, p = np.meshgrid(ct, ph, indexing="ij")
    s = np.sqrt(1.0 - c * c)
    d = np.stack([s * np.cos(p), s * np.sin(p), c], -1).reshape(-1, 3)
    return d.astype(np.float32), (1.0 / nt) * (2.0 * np.pi / nphi)


def u_grid(n=400):
    g = (np.arange(n) + 0.5) / n
    u, v = np.meshgrid(g, g, indexing="ij")
    return np.stack([u.ravel(), v.ravel()], -1)


def eval_rows(m, wo, dirs):
    q = np.zeros((len(dirs), 12), np.float32)
    q[:, 0:3], q[:, 3:6], q[:, 6:9] = wo, dirs, Z
    return O.bsdf_query([m], np.zeros(len(q), np.int32), q)


def sample_rows(m, wo, u):
    q = np.zeros((len(u), 12), np.float32)
    q[:, 0:3], q[:, 6:9], q[:, 9:11] = wo, Z, u
    return O.bsdf_query([m], np.zeros(len(q), np.int32), q)


WOS = [B.unit([0.3, 0.2, 0.9]), B.unit([0.8, -0.1, 0.3])]


@pytest.mark.parametrize("wo", WOS)
def test_lambert_closed_forms(wo):
    m = B.MATERIALS[0]
    d, dw = sphere_grid()
    r = eval_rows(m, wo, d)
    assert abs(r[:, 3].sum() * dw - 1.0) < 2e-4
    assert np.allclose((r[:, 0:3] * np.abs(d[:, 2:3])).sum(0) * dw, m["reflectance"], rtol=2e-4)
    s = sample_rows(m, wo, u_grid(64))
    assert np.all(s[:, 11] == 1) and np.allclose(s[:, 7:10], m["reflectance"])
    assert np.allclose(s[:, 10], s[:, 6] / np.float32(np.pi), rtol=1e-6)  # pdf = cos / pi


@pytest.mark.parametrize("mat", [2, 4])
@pytest.mark.parametrize("wo", WOS)
def test_ggx_conductor_pdf_and_estimator(mat, wo):
    m = B.MATERIALS[mat]
    d, dw = sphere_grid()
    r = eval_rows(m, wo, d)
    s = sample_rows(m, wo, u_grid())
    valid = (s[:, 11].astype(int) & 1) == 1
    assert not ((s[:, 11].astype(int) & 2) == 2).any()
    # the pdf is the density of the sampler: it integrates to P(valid sample)
    assert abs(r[:, 3].sum() * dw - valid.mean()) < 2e-3
    # and the estimator's mean is the integral of f cos
    quad = (r[:, 0:3] * np.abs(d[:, 2:3])).sum(0) * dw
    est = (s[:, 7:10] * valid[:, None]).mean(0)
    assert np.allclose(est, quad, rtol=3e-3, atol=1e-3), (est, quad)


@pytest.mark.parametrize("mat", [2, 4, 6, 7])
def test_reflection_is_reciprocal(mat):
    m = B.MATERIALS[mat]
    rng = np.random.default_rng(5)
    n = 20000
    wo, wi = B.unit(rng.normal(size=(n, 3))), B.unit(rng.normal(size=(n, 3)))
    same = wo[:, 2] * wi[:, 2] > 0
    wo, wi = wo[same], wi[same]
    a = eval_rows(m, wo, wi)
    q = np.zeros((len(wo), 12), np.float32)
    q[:, 0:3], q[:, 3:6], q[:, 6:9] = wi, wo, Z
    b = O.bsdf_query([m], np.zeros(len(q), np.int32), q)
    big = a[:, 0] > 1e-4
    assert big.sum() > 100
    assert np.allclose(a[big, 0:3], b[big, 0:3], rtol=2e-5)
    assert np.all((a[:, 0] > 0) == (b[:, 0] > 0))


@pytest.mark.parametrize("mat", [0, 2, 4, 6, 7])
def test_sample_pdf_and_weight_agree_with_evaluation(mat):
    m = B.MATERIALS[mat]
    rng = np.random.default_rng(9)
    for wo in WOS + [B.unit([0.2, 0.1, -0.9])]:
        s = sample_rows(m, wo, rng.uniform(size=(4000, 2)))
        ok = (s[:, 11].astype(int) & 3) == 1
        if not ok.any():
            continue
        r = eval_rows(m, wo, s[ok, 4:7])
        assert np.allclose(r[:, 3], s[ok, 10], rtol=1e-6)
        expect = r[:, 0:3] * np.abs(s[ok, 6:7]) / r[:, 3:4]
        assert np.allclose(s[ok, 7:10], expect, rtol=2e-5, atol=1e-7)


def fresnel(ci, ei, et):
    if ci <= 0:
        ei, et, ci = et, ei, -ci
    st = ei / et * np.sqrt(max(0.0, 1 - ci * ci))
    if st >= 1:
        return 1.0
    ct = np.sqrt(max(0.0, 1 - st * st))
    rpa = (et * ci - ei * ct) / (et * ci + ei * ct)
    rpe = (ei * ci - et * ct) / (ei * ci + et * ct)
    return 0.5 * (rpa * rpa + rpe * rpe)


@pytest.mark.parametrize("mat", [5, 9])
@pytest.mark.parametrize("wo", WOS + [B.unit([0.2, 0.1, -0.9])])
def test_smooth_dielectrics_pick_by_fresnel(mat, wo):
    m = B.MATERIALS[mat]
    u = u_grid(200)
    s = sample_rows(m, wo, u)
    assert np.all(s[:, 11] == 7)  # valid, specular, delta
    assert np.allclose(s[:, 7:10], m["reflectance"]) and np.all(s[:, 10] == 1.0)
    reflect = s[:, 6] * wo[2] > 0
    inside = wo[2] < 0
    ei, et = (1.0, 1.5) if (mat == 9 or not inside) else (1.5, 1.0)
    f = fresnel(abs(wo[2]), ei, et)
    assert abs(reflect.mean() - f) < 1e-2 / 2
    if mat == 9:  # thin: straight through otherwise
        assert np.allclose(s[~reflect, 4:7], -wo, atol=1e-6)


@pytest.mark.parametrize("wo", WOS)
def test_rough_thin_dielectric_transmission_pdf(wo):
    m = B.MATERIALS[10]
    d, dw = sphere_grid()
    r = eval_rows(m, wo, d)
    assert abs(r[:, 3].sum() * dw - 1.0) < 2e-4  # cosine lobe about -wo
