"""CPU pins of the path-tracing oracle.

* the reference's known answers (tests/backends/cpu/cpu_path_tracer_lighting_tests.cpp:6-18,
  cpu_path_tracer_execution_tests.cpp:10-35);
* a closed form: a white-furnace-style constant environment seen through one
  bounce of a diffuse triangle;
* bit-identical films and counters against the reference's own
  RenderCpuPathTrace (oracle/_ref), for the Cornell workload and a two-level
  environment-lit scene, across all four samplers.
"""
import numpy as np
import pytest

from oracle import oracle_py as O
from tests.helpers import base_path_scene, material_zoo_scene
from yaoray_amd import Scene

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (reference sources absent)")


def linear(film, w, h):
    return (film["sum"] / np.maximum(film["samples"], 1)[:, None]).reshape(h, w, 3)


def test_emissive_surface_known_answer():
    s = base_path_scene()
    s.set_materials([{"kind": 0, "reflectance": (0.0, 0.0, 0.0), "emission": (0.25, 0.5, 0.75)}])
    s.set_constant_environment((0.0, 0.0, 0.0))
    film, _ = O.render(s, 3, 3, 1, max_depth=1, sampler=0, seed=7)
    c = linear(film, 3, 3)[1, 1]
    assert np.allclose(c, [0.25, 0.5, 0.75], atol=1e-6)


def test_ray_counts_known_answers():
    s = base_path_scene()
    film, st = O.render(s, 4, 3, 1, max_depth=1, sampler=0, seed=7)
    assert st["rays_traced"] == 12 and st["hits"] + st["misses"] == 12
    assert np.all(film["samples"] == 1)
    film, st = O.render(s, 2, 2, 4, max_depth=1, sampler=0, seed=7)
    assert st["rays_traced"] == 16 and np.all(film["samples"] == 4)


def test_miss_returns_environment_and_one_bounce_closed_form():
    # Rays that miss see the constant environment; with max_depth 2 a camera
    # hit on the diffuse triangle (albedo 0.8) adds L_env * 0.8 * (fraction of
    # its hemisphere sampling directions that escape = all, since nothing else
    # is in the scene) in expectation, i.e. exactly 0.8 * L_env per sample
    # because the cosine-weighted estimator has weight rho.
    s = base_path_scene()
    env = np.array([0.02, 0.03, 0.04], np.float32)
    film, _ = O.render(s, 9, 9, 4, max_depth=2, sampler=2, seed=3)
    img = linear(film, 9, 9)
    assert np.allclose(img[0, 0], env, atol=1e-7)  # corner pixel misses
    assert np.allclose(img[4, 4], 0.8 * env, rtol=1e-5)


@needs_ref
@pytest.mark.parametrize("sampler", [0, 1, 2, 3])
def test_cornell_matches_reference_bits(sampler):
    s = Scene.cornell(1)
    s.build_accel()
    a, sa = O.render(s, 24, 24, 4, max_depth=8, sampler=sampler)
    b, sb, _ = O.ref_render(s, 24, 24, 4, max_depth=8, sampler=sampler)
    assert a.tobytes() == b.tobytes()
    assert sa == {k: sb[k] for k in sa}


@needs_ref
def test_two_level_environment_scene_matches_reference_bits():
    s = Scene.instanced(instances=150, subdivisions=2, variants=6, env_width=64, env_height=32, seed=4)
    s.build_accel()
    for depth, spp in ((1, 1), (12, 3)):
        a, sa = O.render(s, 40, 24, spp, max_depth=depth)
        b, sb, _ = O.ref_render(s, 40, 24, spp, max_depth=depth)
        assert a.tobytes() == b.tobytes()
        assert sa == {k: sb[k] for k in sa}


@needs_ref
def test_known_answer_scene_matches_reference_bits():
    s = base_path_scene()
    a, sa = O.render(s, 8, 6, 2, max_depth=3, sampler=1, seed=7)
    b, sb, _ = O.ref_render(s, 8, 6, 2, max_depth=3, sampler=1, seed=7)
    assert a.tobytes() == b.tobytes() and sa == {k: sb[k] for k in sa}


@needs_ref
def test_material_zoo_scene_matches_reference_bits():
    """Every in-scope BSDF branch in a render: rough and smooth dielectrics (one
    absorbing), a delta mirror, smooth and rough thin dielectrics, diffuse and
    rough conductor instances under the sky environment."""
    s = material_zoo_scene()
    for depth, spp in ((1, 1), (12, 3)):
        a, sa = O.render(s, 64, 36, spp, max_depth=depth)
        b, sb, _ = O.ref_render(s, 64, 36, spp, max_depth=depth)
        assert a.tobytes() == b.tobytes()
        assert sa == {k: sb[k] for k in sa}
