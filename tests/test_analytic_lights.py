"""Analytic (point / spot / distant) lights on the CPU: the oracle's
SampleAnalyticLight and next-event term and the host's selection tables,
pinned by

  * the reference's own known answers (tests/lighting/analytic_light_tests.cpp:8-125,
    tests/lighting/light_sampling_tests.cpp:258-279,
    tests/backends/cpu/cpu_path_tracer_lighting_tests.cpp:51-84);
  * what the mathematics fixes: selection frequencies over a fine grid of
    selection samples equal the returned selection pdfs (alias table and light
    tree alike), the pdfs sum to one, and a depth-1 render of a diffuse sphere
    under one point light equals rho / pi * I cos / d^2 in closed form;
  * the reference itself (oracle/_ref): the power alias table and the light
    tree byte for byte, and bit-identical films and counters of scenes with
    every light kind, with and without the tree; skipped where _ref is absent.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle_py as O
from tests.helpers import (CORNELL_LIGHTS, cornell_lights_scene, lit_sphere_scene, many_point_lights_scene,
                           same_bits)
from yaoray_amd import Scene
from yaoray_amd import native as N

SPOT = {"kind": "spot", "position": (0.0, 1.0, 0.0), "direction": (0.0, -1.0, 0.0), "intensity": (4.0, 4.0, 4.0),
        "cone_angle": float(np.float32(np.cos(np.float32(0.5236)))),
        "cone_cos_inner": float(np.float32(np.cos(np.float32(0.4363))))}


def one_light(light):
    return lit_sphere_scene([light])


def sample(scene, points, u=0.5):
    pts = np.asarray(points, np.float32).reshape(-1, 3)
    return O.sample_analytic(scene, pts, np.full(len(pts), u, np.float32))


# -- reference known answers ---------------------------------------------------------

def test_point_light_known_answers():
    s = one_light({"kind": "point", "position": (0.0, 2.0, 0.0), "intensity": (10.0, 10.0, 10.0)})
    r = sample(s, [(0.0, 0.0, 0.0)])
    assert r["valid"][0] and abs(r["distance"][0] - 2.0) < 1e-5
    assert np.allclose(r["wi"][0], (0.0, 1.0, 0.0), atol=1e-5)
    assert np.allclose(r["radiance"][0], 2.5, atol=1e-5)  # I / d^2
    s = one_light({"kind": "point", "position": (1.0, 2.0, 3.0), "intensity": (1.0, 1.0, 1.0)})
    assert not sample(s, [(1.0, 2.0, 3.0)])["valid"][0]  # zero distance


def test_distant_light_known_answers():
    s = one_light({"kind": "distant", "direction": (0.0, -1.0, 0.0), "intensity": (2.0, 3.0, 4.0)})
    r = sample(s, [(1.0, 2.0, 3.0), (10.0, 20.0, 30.0)])
    assert np.all(r["valid"]) and np.all(r["distance"] > 1e5)
    assert np.allclose(r["wi"], (0.0, 1.0, 0.0), atol=1e-6)
    assert np.allclose(r["radiance"], (2.0, 3.0, 4.0), atol=1e-6)


def test_spot_light_known_answers():
    s = one_light(SPOT)
    inside, outside, band = (sample(s, [p]) for p in ((0.0, 0.0, 0.0), (2.0, 0.0, 0.0), (0.521, 0.0, 0.0)))
    assert inside["valid"][0] and abs(inside["distance"][0] - 1.0) < 1e-5
    assert np.allclose(inside["radiance"][0], 4.0, atol=1e-5)
    assert not outside["valid"][0] or abs(outside["radiance"][0][0]) < 1e-5
    assert band["valid"][0] and 0.0 < band["radiance"][0][0] < 4.0
    # The falloff band is the smoothstep of the cosine between the cone cosines.
    d = np.array([0.0, 1.0, 0.0]) - np.array([0.521, 0.0, 0.0])
    cosine = d[1] / np.linalg.norm(d)
    t = (cosine - SPOT["cone_angle"]) / (SPOT["cone_cos_inner"] - SPOT["cone_angle"])
    assert abs(band["radiance"][0][0] - 4.0 * t * t * (3 - 2 * t) / d.dot(d)) < 1e-4


def test_light_tree_prefers_nearby_lights():
    s = lit_sphere_scene([{"kind": "point", "position": (10.0 * i, 0.0, 0.0), "intensity": (1.0, 1.0, 1.0)}
                          for i in range(12)])
    sd = s.shading_desc()
    assert sd.use_analytic_tree == 1 and sd.analytic_tree_count == 23
    u = (np.arange(4096, dtype=np.float32) + 0.5) / 4096
    r = O.sample_analytic(s, np.tile([0.1, 0.0, 0.0], (4096, 1)), u)
    assert np.all(r["valid"]) and np.all(r["selection_pdf"] > 0)
    counts = np.bincount(r["light_index"], minlength=12)
    assert counts[0] > counts[11]


# -- what the mathematics fixes -------------------------------------------------------

@pytest.mark.parametrize("kind", ["alias", "tree"])
def test_selection_frequencies_equal_selection_pdfs(kind):
    """The selection sample u is uniform, so the measure of u that selects light
    i is its probability: frequencies over a fine u grid equal the returned
    pdfs, which sum to one over the lights."""
    s = cornell_lights_scene(False) if kind == "alias" else many_point_lights_scene(20, seed=8)
    sd = s.shading_desc()
    assert sd.use_analytic_tree == (1 if kind == "tree" else 0)
    n = sd.analytic_light_count
    m = 1 << 18
    u = (np.arange(m, dtype=np.float64) + 0.5) / m
    rng = np.random.default_rng(2)
    for p in rng.uniform((-0.9, 0.1, -0.9), (0.9, 1.9, 0.9), (5, 3)):
        r = O.sample_analytic(s, np.tile(p, (m, 1)), u.astype(np.float32))
        pdf = np.zeros(n)
        for i in range(n):
            sel = r["light_index"] == i
            if sel.any():
                assert np.all(r["selection_pdf"][sel] == r["selection_pdf"][sel][0])
                pdf[i] = r["selection_pdf"][sel][0]
        freq = np.bincount(r["light_index"][r["light_index"] >= 0], minlength=n) / m
        # A light no grid point selects has probability below about 1 / m.
        unseen = int(np.sum(pdf == 0))
        assert -1e-5 < 1.0 - pdf.sum() < 1e-5 + 2.0 * unseen / m
        assert np.allclose(freq, pdf, atol=4.0 / m + 1e-6), (freq, pdf)


def test_point_light_depth1_closed_form():
    """One point light above a diffuse sphere, no other light: a camera ray
    that hits the sphere at x returns rho / pi * I * cos / d^2 (depth 1: the
    next-event term only)."""
    light = (0.0, 2.0, 1.0)
    s = lit_sphere_scene([{"kind": "point", "position": light, "intensity": (40.0, 40.0, 40.0)}])
    w = h = 64
    film, st = O.render(s, w, h, 1, max_depth=1, sampler=0)
    org, d = s.camera_rays(w, h)
    v, i, p = s.geometry()
    hit = O.trace_flat(v, i, p, s.flat_bvh(), org, d, spheres=s.spheres())
    img = film["sum"] / np.maximum(film["samples"], 1)[:, None]
    on = hit["triangle"] == -2
    assert on.sum() > 250 and st["rays_traced"] == w * h
    t = np.where(on, hit["t"], 0.0)
    x = org[:, :3].astype(np.float64) + d[:, :3] * t[:, None]
    n = x / np.maximum(np.linalg.norm(x, axis=1, keepdims=True), 1e-12)
    to = np.array(light) - x
    dist2 = np.einsum("ij,ij->i", to, to)
    cos = np.einsum("ij,ij->i", n, to) / np.sqrt(dist2)
    expect = np.where(cos > 0, 0.8 / np.pi * 40.0 * np.maximum(cos, 0) / dist2, 0.0)
    # The camera ray of the film and of camera_rays is the pixel centre only
    # for sampler 0 at one sample; the sphere self-shadows nothing here.
    lit = on & (cos > 1e-3)
    assert np.allclose(img[lit, 0], expect[lit], rtol=2e-4, atol=1e-6)
    assert np.all(img[on & (cos < -1e-3), 0] == 0.0)
    assert np.all(img[~on] == 0.0)


def test_reference_cpu_tests_pass_on_the_oracle():
    """cpu_path_tracer_lights_a_sphere_with_a_point_light / _distant_light:
    a lit centre pixel, and under a light from above the top of the sphere is
    brighter than its bottom."""
    s = lit_sphere_scene([{"kind": "point", "position": (0.0, 2.0, 1.0), "intensity": (40.0, 40.0, 40.0)}])
    film, st = O.render(s, 32, 32, 8, max_depth=2)
    img = (film["sum"] / np.maximum(film["samples"], 1)[:, None]).reshape(32, 32, 3)
    assert st["hits"] > 0 and np.all(img[16, 16] > 0)
    s = lit_sphere_scene([{"kind": "distant", "direction": (0.0, -1.0, 0.0), "intensity": (5.0, 5.0, 5.0)}])
    film, _ = O.render(s, 32, 32, 8, max_depth=2)
    img = (film["sum"] / np.maximum(film["samples"], 1)[:, None]).reshape(32, 32, 3)
    assert img[13, 16, 0] > img[19, 16, 0]


# -- the reference itself ------------------------------------------------------------

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (reference sources absent)")


def light_array(scene):
    sd = scene.shading_desc()
    n = sd.analytic_light_count
    return (N.AnalyticLight * n).from_buffer_copy(C.string_at(sd.analytic_lights, n * C.sizeof(N.AnalyticLight)))


@needs_ref
@pytest.mark.parametrize("make", [lambda: cornell_lights_scene(False), lambda: many_point_lights_scene(12),
                                  lambda: many_point_lights_scene(37, seed=5)])
def test_host_tables_match_reference(make):
    s = make()
    sd = s.shading_desc()
    n = sd.analytic_light_count
    prob, alias, pmf, tree, root, use = O.ref_analytic_tables(light_array(s))
    def arr(ptr, ctype):
        return np.ctypeslib.as_array((ctype * n).from_address(ptr)).copy()

    assert same_bits(arr(sd.analytic_alias_probability, C.c_float), prob)
    assert np.array_equal(arr(sd.analytic_alias_index, C.c_int32), alias)
    assert same_bits(arr(sd.analytic_alias_pmf, C.c_float), pmf)
    assert sd.use_analytic_tree == use and sd.analytic_tree_root == root
    assert sd.analytic_tree_count == len(tree)
    if len(tree):
        assert C.string_at(sd.analytic_tree, 40 * len(tree)) == tree.tobytes()


@needs_ref
@pytest.mark.parametrize("kind", ["lit_sphere", "cornell", "cornell_spheres", "tree"])
def test_oracle_render_matches_reference_with_analytic_lights(kind):
    if kind == "lit_sphere":
        s = lit_sphere_scene([CORNELL_LIGHTS[0], {"kind": "distant", "direction": (0.0, -1.0, 0.0),
                                                  "intensity": (5.0, 5.0, 5.0)}])
        w, h, spp, depth = 32, 32, 8, 2
    elif kind == "tree":
        s = many_point_lights_scene(12)
        w, h, spp, depth = 40, 32, 4, 6
    else:
        s = cornell_lights_scene(kind == "cornell_spheres")
        w, h, spp, depth = 40, 32, 4, 8
    o, ost = O.render(s, w, h, spp, max_depth=depth)
    r, rst, _ = O.ref_render(s, w, h, spp, max_depth=depth)
    for f in o.dtype.names:
        assert same_bits(o[f], r[f]), f
    for k, val in ost.items():
        assert rst[k] == val, k
