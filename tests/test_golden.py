"""CPU pins against the golden fixtures (the reference's own outputs, written
by tests/golden/make_golden.py from oracle/_ref): they hold where the
reference is absent, so the oracle and the host builder stay pinned on any
machine.

* yr::Sampler values, all four kinds, three sample counts, 1D and 2D;
* SAH node arrays of the host builder, byte for byte (flat and TLAS);
* per-ray hits and node / triangle counters of the oracle, bit for bit;
* the Cornell workload's FP32 film and statistics from the oracle, bit for bit;
* analytic spheres: per-ray hits (t, SphereUv, id) and node / triangle /
  sphere counters of the oracle on flat and TLAS spheres, and the sphere
  Cornell box's FP32 film and statistics, bit for bit;
* the reference's own Cornell scene (scenes/pbrt/cornell_box_pbrt: mirror and
  glass spheres, an area light and a point light) as its PBRT frontend
  compiles it: the oracle's film and statistics equal its CPU backend's.
"""
import numpy as np
import pytest

from oracle import oracle_py as O
from tests import golden_data as G
from tests.helpers import same_bits

TRACE_KEYS = ("t", "u", "v", "triangle", "node_tests", "triangle_tests")


@pytest.mark.parametrize("kind", [0, 1, 2, 3])
def test_oracle_sampler_matches_fixture(kind):
    f = G.load("sampler")
    seed = int(f["seed"])
    for spp in (1, 13, 16):
        s = np.minimum(f["sample"], spp - 1)
        o1 = O.sample_values(kind, seed, spp, f["x"], f["y"], s, f["dim"], 1)
        o2 = O.sample_values(kind, seed, spp, f["x"], f["y"], s, f["dim"], 2)
        assert same_bits(o1, f[f"v1_{kind}_{spp}"]), spp
        assert same_bits(o2, f[f"v2_{kind}_{spp}"]), spp


def test_builder_and_oracle_match_flat_fixture():
    f = G.load("flat_trace")
    s = G.flat_scene(f)
    bvh = s.flat_bvh()
    assert bvh.nodes.tobytes() == G.nodes(f["nodes"]).tobytes()
    assert np.array_equal(bvh.primitive_indices, f["primitive_indices"])
    assert bvh.max_depth == int(f["max_depth"])
    v, i, p = s.geometry()
    o = O.trace_flat(v, i, p, bvh, f["org"], f["dir"])
    exp = G.expected_hits(f)
    for k in TRACE_KEYS:
        assert same_bits(o[k], exp[k]), k
    assert 0.2 < (exp["triangle"] >= 0).mean() < 0.9


def test_builder_and_oracle_match_two_level_fixture():
    f = G.load("two_level_trace")
    s = G.two_level_scene(f)
    tl, tp, inst = s.tlas()
    assert tl.nodes.tobytes() == G.nodes(f["tlas_nodes"]).tobytes()
    assert np.array_equal(tl.primitive_indices, f["tlas_primitive_indices"])
    v, i, p = s.geometry()
    blases = [s.blas(k) for k in range(s.info.blas_count)]
    o = O.trace_two_level(v, i, p, blases, inst, tl, tp, f["org"], f["dir"])
    exp = G.expected_hits(f)
    for k in TRACE_KEYS + ("instance",):
        assert same_bits(o[k], exp[k]), k


@pytest.mark.parametrize("depth", [1, 8])
def test_oracle_render_matches_cornell_fixture(depth):
    f = G.load("cornell_film")
    s = G.cornell_scene(f)
    w, h = int(f["width"]), int(f["height"])
    spp = int(f[f"spp{depth}"])
    film, st = O.render(s, w, h, spp, max_depth=depth, sampler=2, seed=int(f["seed"]))
    assert film.tobytes() == G.film(f[f"film{depth}"]).tobytes()
    assert st == G.stats(f[f"stats{depth}"])


@pytest.mark.parametrize("depth", [1, 8])
def test_oracle_adaptive_render_matches_reference_fixture(depth):
    """The restated active-pixel loop with Film::IsConverged against the
    reference's adaptive RenderCpuPathTrace: per-pixel sample counts, film and
    counters bit for bit, and the converged-pixel count."""
    f = G.load("cornell_adaptive")
    s = G.cornell_scene(f)
    w, h = int(f["width"]), int(f["height"])
    spp = int(f[f"spp{depth}"])
    a = tuple(float(v) for v in f[f"adaptive{depth}"])
    film, st, passes, conv = O.render_adaptive(s, w, h, spp, (int(a[0]), *a[1:]), max_depth=depth, sampler=2,
                                               seed=int(f["seed"]))
    exp = G.film(f[f"film{depth}"])
    assert np.array_equal(film["samples"], exp["samples"])
    assert film.tobytes() == exp.tobytes()
    assert st == G.stats(f[f"stats{depth}"])
    assert conv == int(f[f"converged{depth}"])
    assert 0 < conv < w * h and passes == spp  # some pixels converge early, some run to spp


def test_oracle_adaptive_limits():
    """Closed cases of the convergence test: with adaptive sampling off the
    loop is the plain pass loop; a tolerance no pixel can miss stops every
    pixel at max(2, min_spp) samples; a zero tolerance stops only pixels
    whose luminance never varied."""
    s = G.cornell_scene(G.load("cornell_film"))
    plain, pst = O.render(s, 24, 24, 6, max_depth=4)
    off, ost, passes, conv = O.render_adaptive(s, 24, 24, 6, None, max_depth=4)
    assert off.tobytes() == plain.tobytes() and ost == pst and passes == 6 and conv == 0
    loose, _, passes, conv = O.render_adaptive(s, 24, 24, 12, (3, 1e30, 1e30, 1.96), max_depth=4)
    assert np.all(loose["samples"] == 3) and passes == 3 and conv == 24 * 24
    tight, _, _, conv = O.render_adaptive(s, 24, 24, 8, (2, 0.0, 0.0, 1.96), max_depth=4)
    still = tight["luminance_m2"] == 0.0
    assert np.all(tight["samples"][still] == 2) and np.all(tight["samples"][~still] == 8)
    assert conv == int(still.sum())


def test_oracle_render_matches_zoo_fixture():
    """Every in-scope BSDF branch (rough / smooth / absorbing / thin dielectrics,
    mirror and rough conductors, diffuse) under the sky environment: the oracle's
    FP32 film and counters equal the reference's bit for bit."""
    f = G.load("zoo_film")
    s = G.zoo_scene(f)
    w, h = int(f["width"]), int(f["height"])
    film, st = O.render(s, w, h, int(f["spp"]), max_depth=int(f["depth"]), sampler=2, seed=int(f["seed"]))
    assert film.tobytes() == G.film(f["film"]).tobytes()
    assert st == G.stats(f["stats"])


def test_oracle_matches_spheres_fixture():
    f = G.load("spheres")
    c, z = G.sphere_scenes(f)
    v, i, p = c.geometry()
    o = O.trace_flat(v, i, p, c.flat_bvh(), f["flat_org"], f["flat_dir"], spheres=c.spheres())
    exp = G.expected_hits(f, "flat_")
    for k in TRACE_KEYS + ("sphere_tests",):
        assert same_bits(o[k], exp[k]), k
    assert (exp["triangle"] <= -2).sum() > 500
    v, i, p = z.geometry()
    tl, tp, inst = z.tlas()
    o = O.trace_two_level(v, i, p, [z.blas(k) for k in range(z.info.blas_count)], inst, tl, tp, f["tl_org"],
                          f["tl_dir"], spheres=z.spheres())
    exp = G.expected_hits(f, "tl_")
    for k in TRACE_KEYS + ("instance", "sphere_tests"):
        assert same_bits(o[k], exp[k]), k
    assert (exp["triangle"] <= -2).sum() > 300
    w, h = int(f["width"]), int(f["height"])
    film, st = O.render(c, w, h, int(f["spp"]), max_depth=int(f["depth"]), sampler=2, seed=int(f["seed"]))
    assert film.tobytes() == G.film(f["film"]).tobytes()
    assert st == G.stats(f["stats"])


def test_oracle_matches_reference_cornell_pbrt_fixture():
    f = G.load("cornell_pbrt")
    for k in f:
        if k.startswith("full_") and k != "full_settings":
            assert np.array_equal(f[k], f["smoke_" + k[5:]]), k  # the smoke variant is the same scene
    s, st = G.pbrt_scene(f)
    assert s.info.sphere_count == 2 and s.info.analytic_light_count == 1 and s.info.emissive_count == 1
    film, stats = O.render(s, st["width"], st["height"], int(f["spp"]), max_depth=st["max_depth"],
                           sampler=st["sampler"], seed=st["seed"])
    assert film.tobytes() == G.film(f["film"]).tobytes()
    assert stats == G.stats(f["stats"])


@pytest.mark.parametrize("tag", ["ours", "tt"])
def test_oracle_matches_reference_textures_pbrt_fixture(tag):
    """Textured PBRT scenes through the reference's frontend: this repository's
    hip_textures.pbrt and the reference's texture_test smoke scene; the
    oracle's film and counters equal the reference CPU backend's bit for bit."""
    f = G.load("textures_pbrt")
    s, st = G.pbrt_scene(f, tag)
    assert s.info.texture_count >= 5
    film, stats = O.render(s, st["width"], st["height"], int(f["spp"]), max_depth=st["max_depth"],
                           sampler=st["sampler"], seed=st["seed"])
    assert film.tobytes() == G.film(f[tag + "_film"]).tobytes()
    assert stats == G.stats(f[tag + "_stats"])
