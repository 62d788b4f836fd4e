"""Analytic spheres (yr::RenderSphere) on the CPU: the oracle's sphere test and
the host builder, pinned by

  * the reference's own known answers (tests/geometry/sphere_tests.cpp:6-71,
    tests/accel/bvh_traversal_tests.cpp:47-89, tests/accel/bvh_builder_tests.cpp:46-62
    and :156-164, tests/accel/two_level_bvh_tests.cpp:110-123);
  * the closed form of a ray-sphere intersection in float64 (t, SphereUv's
    longitude/latitude), brute force over every primitive (tree-independent);
  * the reference itself (oracle/_ref): byte-identical trees with spheres in the
    flat BVH and the TLAS, bit-identical hits (t, uv, id) and counters, and a
    bit-identical film of a sphere scene; skipped where _ref is absent.

Hit ids: -1 miss, >= 0 triangle, -2 - s sphere s (the C ABI's encoding)."""
import numpy as np
import pytest

from oracle import oracle_py as O
from tests.helpers import MeshArrays, rays, same_bits, sphere_cornell, sphere_zoo
from yaoray_amd import Scene, YrhError
from yaoray_amd.api import SPHERE_DTYPE


def sphere_array(*spheres):
    a = np.zeros(len(spheres), SPHERE_DTYPE)
    for k, sp in enumerate(spheres):
        a[k]["center"] = sp[0]
        a[k]["radius"] = sp[1]
        a[k]["material_index"] = sp[2] if len(sp) > 2 else 0
        a[k]["area_light_index"] = -1
    return a


def one_sphere(center=(0.0, 0.0, 0.0), radius=1.0, t_min=1e-5, t_max=1e6, org=(0, 0, 3), d=(0, 0, -1)):
    sp = sphere_array((center, radius))
    m = MeshArrays()
    v, i, p = m.arrays()
    o, dd = rays([org], [d], t_min=t_min, t_max=np.float32(t_max))
    return O.brute_force(v, i, p, o, dd, spheres=sp)


# -- reference known answers --------------------------------------------------------

def test_sphere_known_answers():
    h = one_sphere()
    assert h["triangle"][0] == -2 and abs(h["t"][0] - 2.0) < 1e-5       # hits from outside
    assert one_sphere(org=(2, 0, 3))["triangle"][0] == -1                # passes outside
    h = one_sphere(org=(0, 0, 0), d=(0, 0, 1))
    assert h["triangle"][0] == -2 and abs(h["t"][0] - 1.0) < 1e-5       # nearest root from inside
    assert one_sphere(t_max=1.5)["triangle"][0] == -1                    # respects t_max
    assert one_sphere(org=(0, 0, 0), d=(0, 0, 1), t_min=2.0)["triangle"][0] == -1  # respects t_min


def test_sphere_uv_known_answers():
    # SphereUv: north pole v = 1, south pole v = 0, +z front seam u = 0.5.
    north = one_sphere(org=(0, 3, 0), d=(0, -1, 0))
    south = one_sphere(org=(0, -3, 0), d=(0, 1, 0))
    front = one_sphere()
    assert abs(north["v"][0] - 1.0) < 1e-5 and abs(south["v"][0]) < 1e-5
    assert abs(front["u"][0] - 0.5) < 1e-5


def test_flat_bvh_with_spheres_known_answers():
    # A sphere without any triangle: one leaf, one sphere test, node tests > 0.
    s = MeshArrays().scene()
    s.set_spheres([((0, 0, 0), 1.0, 0)])
    s.build_accel()
    b = s.flat_bvh()
    assert len(b.nodes) == 1 and len(b.primitive_indices) == 1 and len(b.primitives) == 1
    assert b.nodes[0]["primitive_count"] == 1
    assert b.primitives[0]["kind"] == 1 and b.primitives[0]["sphere"] == 0
    v, i, p = s.geometry()
    o, d = rays([(0, 0, 3)], [(0, 0, -1)])
    h = O.trace_flat(v, i, p, b, o, d, spheres=s.spheres())
    assert h["triangle"][0] == -2 and abs(h["t"][0] - 2.0) < 1e-5
    assert h["node_tests"][0] > 0 and h["sphere_tests"][0] == 1 and h["triangle_tests"][0] == 0
    # The closer sphere wins over a triangle behind it.
    m = MeshArrays().add_test_triangle(0.0, z=-1.0)
    s2 = m.scene()
    s2.set_spheres([((0, 0, 0), 0.5, 0)])
    s2.build_accel()
    v, i, p = s2.geometry()
    h = O.trace_flat(v, i, p, s2.flat_bvh(), o, d, spheres=s2.spheres())
    assert h["triangle"][0] == -2 and abs(h["t"][0] - 2.5) < 1e-5


def test_two_level_keeps_spheres_in_the_tlas():
    m = MeshArrays().add_test_triangle(0.0, z=-2.0)
    s = m.scene(np.zeros(1, np.int32), np.eye(4, dtype=np.float32).reshape(1, 16))
    s.set_spheres([((0, 0, 0), 0.5, 0)])
    s.build_accel()
    assert s.info.accel_kind == 1
    v, i, p = s.geometry()
    tl, tp, inst = s.tlas()
    assert sorted(tp["kind"].tolist()) == [0, 1]
    o, d = rays([(0, 0, 1)], [(0, 0, -1)])
    h = O.trace_two_level(v, i, p, [s.blas(0)], inst, tl, tp, o, d, spheres=s.spheres())
    assert h["triangle"][0] == -2 and h["instance"][0] == -1 and abs(h["t"][0] - 0.5) < 1e-6


@pytest.mark.parametrize("radius", [0.0, -1.0, np.nan, np.inf])
def test_builder_rejects_invalid_sphere(radius):
    s = MeshArrays().scene()
    s.set_spheres([((0, 0, 0), radius, 0)])
    with pytest.raises(YrhError, match="invalid sphere data"):
        s.build_accel()


# -- closed form and brute force ----------------------------------------------------

def test_sphere_hits_match_the_closed_form():
    rng = np.random.default_rng(11)
    n = 400
    c = rng.uniform(-2, 2, (n, 3))
    r = rng.uniform(0.2, 2.0, n)
    org = rng.uniform(-6, 6, (n, 3))
    aim = c + rng.uniform(-1, 1, (n, 3)) * r[:, None]
    d = aim - org
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    v, i, p = MeshArrays().arrays()
    hits = []
    m = 400  # one sphere per ray
    for k in range(m):
        o4, d4 = rays([org[k]], [d[k]])
        hits.append(O.brute_force(v, i, p, o4, d4, spheres=sphere_array((c[k], r[k])), threads=1))
    t = np.array([h["t"][0] for h in hits])
    ids = np.array([h["triangle"][0] for h in hits])
    u = np.array([h["u"][0] for h in hits])
    vv = np.array([h["v"][0] for h in hits])
    # float64 roots of |o + t d - c|^2 = r^2; the nearer root above t_min, else the farther.
    oc = org[:m] - c[:m]
    b = np.einsum("ij,ij->i", oc, d[:m])
    disc = b * b - (np.einsum("ij,ij->i", oc, oc) - r[:m] ** 2)
    ok = disc >= 0
    sq = np.sqrt(np.maximum(disc, 0))
    t0, t1 = -b - sq, -b + sq
    t_exp = np.where(t0 > 1e-5, t0, np.where(t1 > 1e-5, t1, np.inf))
    hit_exp = ok & np.isfinite(t_exp)
    margin = np.abs(disc) > 1e-3 * r[:m] ** 2  # skip grazing rays
    assert np.array_equal((ids == -2)[margin], hit_exp[margin])
    sel = hit_exp & margin & (ids == -2)
    assert sel.sum() > 100
    assert np.allclose(t[sel], t_exp[sel], rtol=2e-4, atol=2e-4)
    # SphereUv: u = atan2(n.x, n.z) / 2pi + 1/2 (the +z seam at u = 1/2), v = 1 - acos(n.y) / pi.
    nrm = org[:m] + t_exp[:, None] * d[:m] - c[:m]
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    lat = 1.0 - np.arccos(np.clip(nrm[:, 1], -1, 1)) / np.pi
    assert np.allclose(vv[sel], lat[sel], atol=2e-3)
    pole = np.abs(nrm[:, 1]) < 0.99
    lon = np.arctan2(nrm[:, 0], nrm[:, 2]) / (2 * np.pi) + 0.5
    du = np.abs(u - lon)
    du = np.minimum(du, 1 - du)
    assert np.all(du[sel & pole] < 2e-3)


@pytest.mark.parametrize("kind", ["flat", "two_level"])
def test_tree_traversal_equals_brute_force_with_spheres(kind):
    s = sphere_cornell() if kind == "flat" else sphere_zoo(40)
    v, i, p = s.geometry()
    org, d = s.camera_rays(96, 64)
    sp = s.spheres()
    if kind == "flat":
        h = O.trace_flat(v, i, p, s.flat_bvh(), org, d, spheres=sp)
        bf = O.brute_force(v, i, p, org, d, spheres=sp)
        agree = h["triangle"] == bf["triangle"]
        assert agree.mean() > 0.999
        assert np.all(h["t"][~agree] == bf["t"][~agree])
    else:
        tl, tp, inst = s.tlas()
        h = O.trace_two_level(v, i, p, [s.blas(k) for k in range(s.info.blas_count)], inst, tl, tp, org, d,
                              spheres=sp)
    # Every sphere is seen by the camera, and sphere tests are counted.
    seen = set((-2 - h["triangle"][h["triangle"] <= -2]).tolist())
    assert seen == set(range(len(sp))), seen
    assert h["sphere_tests"].sum() > 0


# -- the reference itself ------------------------------------------------------------

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (reference sources absent)")


@needs_ref
def test_builder_matches_reference_bytes_with_spheres():
    s = sphere_cornell()
    v, i, p = s.geometry()
    ref = O.RefScene(v, i, p, spheres=s.spheres()).bvh()
    mine = s.flat_bvh()
    assert mine.nodes.tobytes() == ref.nodes.tobytes()
    assert mine.primitive_indices.tobytes() == ref.primitive_indices.tobytes()
    assert mine.primitives.tobytes() == ref.primitives.tobytes()
    # Many spheres, mixed with many triangles (SAH binning over both kinds).
    rng = np.random.default_rng(21)
    t = Scene.triangles(5000, seed=3)
    t.set_spheres([(tuple(rng.uniform(-1, 1, 3)), float(rng.uniform(0.005, 0.05)), 0) for _ in range(800)])
    t.build_accel()
    v, i, p = t.geometry()
    ref = O.RefScene(v, i, p, spheres=t.spheres()).bvh()
    assert t.flat_bvh().nodes.tobytes() == ref.nodes.tobytes()
    assert t.flat_bvh().primitives.tobytes() == ref.primitives.tobytes()


@needs_ref
@pytest.mark.parametrize("kind", ["flat", "two_level"])
def test_oracle_matches_reference_with_spheres(kind):
    s = sphere_cornell() if kind == "flat" else sphere_zoo(60)
    v, i, p = s.geometry()
    sp = s.spheres()
    org1, d1 = s.camera_rays(128, 96)
    rng = np.random.default_rng(5)
    n = 20000
    lo, hi = ((-1, 0, -1), (1, 2, 1)) if kind == "flat" else ((-12, 0, 5), (12, 6, 40))
    org2 = np.zeros((n, 4), np.float32)
    org2[:, :3] = rng.uniform(lo, hi, (n, 3))
    org2[:, 3] = 1e-5
    dd = rng.normal(size=(n, 3))
    d2 = np.zeros((n, 4), np.float32)
    d2[:, :3] = dd / np.linalg.norm(dd, axis=1, keepdims=True)
    d2[:, 3] = np.inf
    org, d = np.concatenate([org1, org2]), np.concatenate([d1, d2])
    if kind == "flat":
        ref = O.RefScene(v, i, p, spheres=sp)
        h = O.trace_flat(v, i, p, s.flat_bvh(), org, d, spheres=sp)
        keys = ("t", "u", "v", "triangle", "node_tests", "triangle_tests", "sphere_tests")
    else:
        tl, tp, inst = s.tlas()
        ref = O.RefScene(v, i, p, inst["primitive"].copy(), inst["object_to_world"].copy(), spheres=sp)
        rn, rpi, rprims, _ = ref.tlas()
        assert rn.tobytes() == tl.nodes.tobytes() and rpi.tobytes() == tl.primitive_indices.tobytes()
        h = O.trace_two_level(v, i, p, [s.blas(k) for k in range(s.info.blas_count)], inst, tl, tp, org, d,
                              spheres=sp)
        keys = ("t", "u", "v", "triangle", "instance", "node_tests", "triangle_tests", "sphere_tests")
    r = ref.trace(org, d)
    assert (h["triangle"] <= -2).sum() > 1000
    for k in keys:
        assert same_bits(h[k], r[k]), k


@needs_ref
def test_oracle_render_matches_reference_with_spheres():
    """Full path tracing through sphere hits (material from the sphere, outward
    or flipped normal, no sphere emission, media inside a glass sphere):
    the oracle's film and counters equal the reference's bit for bit."""
    s = sphere_cornell()
    o, ost = O.render(s, 48, 36, 4, max_depth=8)
    r, rst, _ = O.ref_render(s, 48, 36, 4, max_depth=8)
    for f in o.dtype.names:
        assert same_bits(o[f], r[f]), f
    for k, val in ost.items():
        assert rst[k] == val, k
    assert ost["sphere_tests"] > 0
