"""BSDF query cases shared by the CPU pins, the golden fixture writer and the
GPU tests.

A query row is 12 float32: wo (3), wi (3), shading normal (3), u (2), unused.
A result row is 12 float32: EvaluateBsdf (3), PdfBsdf, then SampleBsdf's wi (3),
weight (3), pdf and flags = valid + 2 specular + 4 IsDeltaBsdf
(src/shading/bsdf.cpp:17-329).

MATERIALS covers every branch of the in-scope kinds: Lambert; conductor as a
delta mirror (roughness 0), inside the delta band (0 < r <= 1e-3, where
IsDeltaBsdf and the evaluation disagree, bsdf.cpp:13 vs :312-319) and rough
GGX; smooth and rough dielectrics (two roughnesses, tinted, an IOR below 1
that the reference clamps to 1); smooth and rough thin dielectrics.
"""
import numpy as np

DIFFUSE, CONDUCTOR, DIELECTRIC, THIN = 0, 1, 2, 3

MATERIALS = [
    {"kind": DIFFUSE, "reflectance": (0.6, 0.3, 0.15)},
    {"kind": CONDUCTOR, "reflectance": (1.0, 0.72, 0.32), "uroughness": 0.0},
    {"kind": CONDUCTOR, "reflectance": (1.0, 0.72, 0.32), "uroughness": 0.35},
    {"kind": CONDUCTOR, "reflectance": (0.9, 0.9, 0.9), "uroughness": 5.0e-4},
    {"kind": CONDUCTOR, "reflectance": (0.8, 0.6, 0.9), "uroughness": 0.7},
    {"kind": DIELECTRIC, "reflectance": (1.0, 1.0, 1.0), "ior": 1.5},
    {"kind": DIELECTRIC, "reflectance": (1.0, 1.0, 1.0), "ior": 1.5, "uroughness": 0.35},
    {"kind": DIELECTRIC, "reflectance": (0.9, 0.8, 0.7), "ior": 1.33, "uroughness": 0.6},
    {"kind": DIELECTRIC, "reflectance": (1.0, 1.0, 1.0), "ior": 0.8, "uroughness": 0.2},
    {"kind": THIN, "reflectance": (1.0, 1.0, 1.0), "ior": 1.5},
    {"kind": THIN, "reflectance": (0.9, 0.95, 1.0), "ior": 1.5, "uroughness": 0.3},
]


def unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def queries(n, seed=11):
    """n random query rows per material: normals anywhere on the sphere, wo and
    wi on the whole sphere (so both sides of the surface, transmission and
    total internal reflection occur), u uniform with some exact 0 and 1-ulp
    edges. Returns (material index per row, rows)."""
    rng = np.random.default_rng(seed)
    mats, rows = [], []
    for m in range(len(MATERIALS)):
        nn = unit(rng.normal(size=(n, 3)))
        wo = unit(rng.normal(size=(n, 3)))
        wi = unit(rng.normal(size=(n, 3)))
        # Half the wi near the mirror / straight-through directions, where the
        # glossy lobes are large.
        k = n // 2
        r = wo[:k] - 2.0 * np.sum(wo[:k] * nn[:k], axis=1, keepdims=True) * nn[:k]
        mirror = np.where(rng.uniform(size=(k, 1)) < 0.5, -r, -wo[:k])
        wi[:k] = unit(mirror + 0.3 * rng.normal(size=(k, 3)))
        u = rng.uniform(0.0, 1.0, (n, 2))
        u[:16, 0] = 0.0
        u[16:32, 1] = 0.0
        u[32:48] = np.float32(1.0) - np.float32(2.0 ** -24)
        q = np.zeros((n, 12), np.float32)
        q[:, 0:3], q[:, 3:6], q[:, 6:9], q[:, 9:11] = wo, wi, nn, u
        mats.append(np.full(n, m, np.int32))
        rows.append(q)
    return np.concatenate(mats), np.concatenate(rows)


def query_row(wo, wi, n, u):
    q = np.zeros(12, np.float32)
    q[0:3], q[3:6], q[6:9], q[9:11] = wo, wi, n, u
    return q


# The reference's own known answers (tests/shading/bsdf_diffuse_conductor_tests.cpp:12-116,
# tests/shading/bsdf_dielectric_tests.cpp:10-146), as (name, material, query row).
# Materials follow tests/shading/bsdf_test_materials.hpp.
def _ref_material(kind, roughness=0.0, reflectance=None):
    base = {DIFFUSE: (0.6, 0.3, 0.15), CONDUCTOR: (1.0, 0.72, 0.32), DIELECTRIC: (1.0, 1.0, 1.0),
            THIN: (1.0, 1.0, 1.0)}[kind]
    return {"kind": kind, "reflectance": reflectance or base, "ior": 1.5, "uroughness": roughness}


Z = np.array([0.0, 0.0, 1.0], np.float32)


def known_answer_cases():
    f32 = lambda v: unit(np.asarray(v, np.float32)).astype(np.float32)  # noqa: E731
    return [
        ("diffuse_eval", _ref_material(DIFFUSE), query_row(f32([0, 0.25, 1]), f32([0.5, 0, 1]), Z, (0, 0))),
        ("diffuse_below_wo", _ref_material(DIFFUSE), query_row([0, 0, -1], [0, 0, 1], Z, (0, 0))),
        ("diffuse_below_wi", _ref_material(DIFFUSE), query_row([0, 0, 1], [0, 0, -1], Z, (0, 0))),
        ("diffuse_pdf", _ref_material(DIFFUSE), query_row(Z, f32([0, 1, 1]), Z, (0.25, 0.5))),
        ("polished_conductor", _ref_material(CONDUCTOR), query_row(f32([-0.25, 0, 1]), Z, Z, (0, 0))),
        ("rough_conductor", _ref_material(CONDUCTOR, 0.35),
         query_row(f32([0, 0.25, 1]), f32([0.25, 0, 1]), Z, (0.25, 0.5))),
        ("smooth_conductor_normal", _ref_material(CONDUCTOR), query_row(Z, Z, Z, (0, 0))),
        ("dielectric_normal_incidence", _ref_material(DIELECTRIC), query_row(Z, Z, Z, (0.5, 0.25))),
        ("dielectric_fresnel_reflect", _ref_material(DIELECTRIC, reflectance=(0.8, 0.9, 1.0)),
         query_row(f32([-0.25, 0, 1]), Z, Z, (0.0, 0.25))),
        ("dielectric_tir", _ref_material(DIELECTRIC), query_row(f32([0.9, 0, -0.4358899]), Z, Z, (0.9, 0.25))),
        ("rough_dielectric_reflection", _ref_material(DIELECTRIC, 0.35),
         query_row(f32([0, 0.25, 1]), f32([0.25, 0, 1]), Z, (0.0, 0.35))),
        ("rough_dielectric_transmission", _ref_material(DIELECTRIC, 0.35),
         query_row(f32([0, 0.25, 1]), f32([0.1, -0.05, -1]), Z, (0.0, 0.35))),
        ("rough_dielectric_sample_reflect", _ref_material(DIELECTRIC, 0.35),
         query_row(f32([0, 0.2, 1]), Z, Z, (0.0, 0.35))),
        ("rough_dielectric_sample_transmit", _ref_material(DIELECTRIC, 0.35),
         query_row(f32([0, 0.2, 1]), Z, Z, (0.95, 0.35))),
        ("thin_straight_through", _ref_material(THIN), query_row(f32([0.2, 0, 1]), Z, Z, (0.8, 0.5))),
        ("rough_thin", _ref_material(THIN, 0.3), query_row(f32([0.1, 0, 1]), Z, Z, (0.95, 0.4))),
    ]


def check_known_answers(results):
    """Asserts the reference tests' expectations on result rows keyed by case
    name (each the 12-float row of that case), with their tolerances."""
    r = results
    pi = np.float32(np.pi)
    near = lambda a, b, tol=1e-6: abs(float(a) - float(b)) <= tol  # noqa: E731
    valid = lambda row: int(row[11]) & 1 == 1  # noqa: E731
    spec = lambda row: int(row[11]) & 2 == 2  # noqa: E731
    delta = lambda row: int(row[11]) & 4 == 4  # noqa: E731
    black = lambda row: row[0] == 0 and row[1] == 0 and row[2] == 0  # noqa: E731
    # diffuse (bsdf_diffuse_conductor_tests.cpp:12-73)
    d = r["diffuse_eval"]
    assert near(d[0], 0.6 / pi) and near(d[1], 0.3 / pi) and near(d[2], 0.15 / pi)
    assert black(r["diffuse_below_wo"]) and black(r["diffuse_below_wi"])
    d = r["diffuse_pdf"]
    assert near(d[3], np.float32(1 / np.sqrt(2)) / pi)
    assert valid(d) and not spec(d) and d[6] > 0 and d[10] > 0
    assert near(d[7], 0.6) and near(d[8], 0.3) and near(d[9], 0.15)
    # polished conductor (:75-96): mirror direction, weight = reflectance, pdf 1,
    # no finite f or pdf at the mirror direction (re-queried below by the caller)
    c = r["polished_conductor"]
    assert delta(c) and valid(c) and spec(c)
    assert near(c[4], 0.24253563) and near(c[6], 0.9701425)
    assert near(c[7], 1.0) and near(c[8], 0.72) and near(c[9], 0.32) and near(c[10], 1.0)
    # rough conductor (:98-116)
    c = r["rough_conductor"]
    assert not delta(c) and c[2] > 0 and valid(c) and not spec(c) and c[10] > 0
    assert c[7] >= c[8] >= c[9]
    c = r["smooth_conductor_normal"]
    assert delta(c) and black(c) and near(c[3], 0.0)
    # smooth dielectric (bsdf_dielectric_tests.cpp:10-64)
    e = r["dielectric_normal_incidence"]
    assert delta(e) and valid(e) and spec(e)
    assert near(e[4], 0.0) and near(e[5], 0.0) and near(e[6], -1.0) and near(e[7], 1.0) and near(e[10], 1.0)
    e = r["dielectric_fresnel_reflect"]
    assert valid(e) and spec(e) and e[6] > 0 and near(e[4], 0.24253563)
    assert near(e[7], 0.8) and near(e[8], 0.9) and near(e[9], 1.0)
    e = r["dielectric_tir"]
    assert valid(e) and spec(e) and e[6] < 0 and near(e[10], 1.0)
    # rough dielectric (:66-111)
    e = r["rough_dielectric_reflection"]
    assert not delta(e) and e[2] > 0 and e[3] > 0
    e = r["rough_dielectric_transmission"]
    assert e[2] > 0 and e[3] > 0
    a, b = r["rough_dielectric_sample_reflect"], r["rough_dielectric_sample_transmit"]
    assert valid(a) and valid(b) and not spec(a) and not spec(b)
    assert a[6] > 0 and b[6] < 0 and a[10] > 0 and b[10] > 0
    # thin dielectric (:113-146)
    t = r["thin_straight_through"]
    wo = unit(np.float32([0.2, 0, 1]))
    assert delta(t) and valid(t) and spec(t)
    assert near(t[4], -wo[0]) and near(t[5], -wo[1]) and near(t[6], -wo[2])
    t = r["rough_thin"]
    assert not delta(t) and valid(t) and not spec(t) and t[10] > 0 and t[7] > 0
