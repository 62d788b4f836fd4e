"""GPU parity of the HIP traversal kernels against the CPU oracle, through the C ABI.

Bar: bit-exact hit ids, t, barycentrics and node/triangle counters (integer and
index work on identical FP32 arithmetic), occlusion == closest hit found.
"""
import numpy as np
import pytest

from oracle import oracle_py as O
from tests.helpers import MeshArrays, rays, same_bits, trs
from yaoray_amd import Scene, random_rays

pytestmark = pytest.mark.gpu


def gpu_closest(device, scene, org, d):
    device.upload_scene(scene)
    return device.trace_closest(org, d, counters=True)


def assert_same(g, o, keys=("t", "u", "v", "triangle", "node_tests", "triangle_tests")):
    for k in keys:
        if not same_bits(g[k], o[k]):
            bad = np.nonzero(g[k].view(np.uint32) != o[k].view(np.uint32))[0] if g[k].dtype == np.float32 else \
                np.nonzero(g[k] != o[k])[0]
            raise AssertionError(f"{k}: {len(bad)} mismatches, first at {bad[:5]}: gpu {g[k][bad[:5]]} "
                                 f"oracle {o[k][bad[:5]]}")


@pytest.fixture(scope="module")
def tri_scene():
    s = Scene.triangles(60000, seed=2)
    s.build_accel()
    return s


def workload_rays(scene, n_random=40000, res=160):
    o1, d1 = scene.camera_rays(res, res)
    o2, d2 = random_rays(n_random, seed=3)
    return np.concatenate([o1, o2]), np.concatenate([d1, d2])


def test_flat_closest_matches_oracle(device, tri_scene):
    org, d = workload_rays(tri_scene)
    g = gpu_closest(device, tri_scene, org, d)
    v, i, p = tri_scene.geometry()
    o = O.trace_flat(v, i, p, tri_scene.flat_bvh(), org, d)
    assert_same(g, o)
    assert np.all(g["instance"] == -1)
    assert 0.2 < (g["triangle"] >= 0).mean() < 1.0


def test_flat_any_matches_oracle(device, tri_scene):
    org, d = workload_rays(tri_scene, n_random=20000, res=100)
    device.upload_scene(tri_scene)
    g = device.trace_any(org, d, counters=True)
    v, i, p = tri_scene.geometry()
    o = O.trace_flat(v, i, p, tri_scene.flat_bvh(), org, d, occlusion=True)
    c = O.trace_flat(v, i, p, tri_scene.flat_bvh(), org, d)
    assert np.array_equal(g["occluded"], o["occluded"])
    assert np.array_equal(g["occluded"].astype(bool), c["triangle"] >= 0)
    assert_same(g, o, keys=("node_tests", "triangle_tests"))


def test_finite_segments_and_tmin(device, tri_scene):
    org, d = random_rays(20000, seed=21)
    rng = np.random.default_rng(4)
    d[:, 3] = rng.uniform(0.0, 0.5, len(d)).astype(np.float32)
    org[:, 3] = rng.uniform(0.0, 0.05, len(d)).astype(np.float32)
    g = gpu_closest(device, tri_scene, org, d)
    v, i, p = tri_scene.geometry()
    o = O.trace_flat(v, i, p, tri_scene.flat_bvh(), org, d)
    assert_same(g, o)
    ga = device.trace_any(org, d)
    assert np.array_equal(ga["occluded"].astype(bool), o["triangle"] >= 0)


def test_axis_parallel_and_degenerate_rays(device):
    # Rays with zero direction components (1/d = +-inf, 0*inf = NaN on slab planes),
    # rays lying exactly on box faces, and a zero-length direction.
    m = MeshArrays()
    grid = []
    for x in range(6):
        for y in range(6):
            grid.append([[x, y, 0], [x + 1, y, 0], [x, y + 1, 0]])
            grid.append([[x, y, 1], [x + 1, y, 1], [x, y + 1, 1]])
    m.add_triangles(np.array(grid, np.float32))
    s = m.scene()
    s.build_accel()
    origins, dirs = [], []
    for x in np.linspace(-0.5, 6.5, 15):
        for y in np.linspace(-0.5, 6.5, 15):
            origins += [(x, y, 2.0), (x, y, -1.0), (x, y, 0.0), (x, -1.0, 0.5)]
            dirs += [(0, 0, -1), (0, 0, 1), (1, 0, 0), (0, 1, 0)]
    origins += [(0.2, 0.2, 2.0), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0)]
    dirs += [(0, 0, 0), (-0.0, 0.0, -1.0), (0.0, -0.0, -1.0)]
    org, d = rays(origins, dirs)
    g = gpu_closest(device, s, org, d)
    v, i, p = s.geometry()
    o = O.trace_flat(v, i, p, s.flat_bvh(), org, d)
    assert_same(g, o)


def test_empty_scene_and_single_triangle(device):
    empty = MeshArrays().scene()
    empty.build_accel()
    org, d = rays([(0, 0, 1)] * 3, [(0, 0, -1)] * 3)
    g = gpu_closest(device, empty, org, d)
    assert np.all(g["triangle"] == -1) and np.all(g["node_tests"] == 0)
    assert np.all(g["t"] == np.inf)
    one = MeshArrays().add_test_triangle(0.0).scene()
    one.build_accel()
    g = gpu_closest(device, one, org, d)
    assert np.all(g["triangle"] == 0) and np.all(np.abs(g["t"] - 1.0) < 1e-6)
    assert np.all(g["triangle_tests"] == 1)


def test_deep_tree_uses_stack_spill(device):
    # Triangles at x = 4^k make the SAH peel about one triangle per level, so
    # the tree is deeper than the 16 LDS stack levels; rays travelling along +x
    # in the triangles' plane hit both children at every level and fill the
    # stack to the tree depth.
    tris = []
    for k in range(59):
        x = 4.0 ** k
        tris.append([[x, 0, 0], [x * 1.05, 0, 0], [x, 0.5 * max(1.0, x * 0.01), 0]])
    m = MeshArrays().add_triangles(np.array(tris, np.float32))
    s = m.scene()
    s.build_accel()
    assert s.flat_bvh().max_depth > 20
    xs = np.array([4.0 ** k * 1.01 for k in range(59)] * 4, np.float32)
    ys = np.repeat([0.05, 0.1, 0.2, 0.45], 59).astype(np.float32)
    origins = np.stack([xs, ys, np.full_like(xs, 1.0)], 1)
    org, d = rays(origins, [(0, 0, -1)] * len(xs))
    o2, d2 = rays([(-1.0, 0.1, 0.0)] * 4, [(1, 0, 0), (1, 0.001, 0), (1, 0, 0.001), (1, 0.01, 0)])
    org, d = np.concatenate([org, o2]), np.concatenate([d, d2])
    g = gpu_closest(device, s, org, d)
    v, i, p = s.geometry()
    o = O.trace_flat(v, i, p, s.flat_bvh(), org, d)
    assert_same(g, o)


def test_launch_shape_invariance(device, tri_scene):
    org, d = workload_rays(tri_scene, n_random=30000, res=64)
    base = gpu_closest(device, tri_scene, org, d)
    try:
        for bpc, bs in ((1, 64), (2, 128), (8, 256)):
            device.set_trace_config(bpc, bs)
            assert_same(device.trace_closest(org, d, counters=True), base)
    finally:
        device.set_trace_config(4, 256)


def two_level_scene():
    s = Scene.instanced(instances=500, subdivisions=3, variants=4, env_width=8, env_height=4, seed=4)
    s.build_accel()
    return s


def test_two_level_closest_and_any_match_oracle(device):
    s = two_level_scene()
    v, i, p = s.geometry()
    tl, tp, inst = s.tlas()
    blases = [s.blas(k) for k in range(s.info.blas_count)]
    o1, d1 = s.camera_rays(240, 135)
    rng = np.random.default_rng(5)
    o2 = np.zeros((20000, 4), np.float32)
    o2[:, :3] = rng.uniform([-30, 0, -30], [30, 5, 30], (20000, 3))
    o2[:, 3] = 1e-5
    dd = rng.normal(size=(20000, 3))
    d2 = np.zeros((20000, 4), np.float32)
    d2[:, :3] = dd / np.linalg.norm(dd, axis=1, keepdims=True)
    d2[:, 3] = np.inf
    org, d = np.concatenate([o1, o2]), np.concatenate([d1, d2])
    g = gpu_closest(device, s, org, d)
    o = O.trace_two_level(v, i, p, blases, inst, tl, tp, org, d)
    assert_same(g, o, keys=("t", "u", "v", "triangle", "instance", "node_tests", "triangle_tests"))
    ga = device.trace_any(org, d, counters=True)
    oa = O.trace_two_level(v, i, p, blases, inst, tl, tp, org, d, occlusion=True)
    assert np.array_equal(ga["occluded"], oa["occluded"])
    assert np.array_equal(ga["occluded"].astype(bool), o["triangle"] >= 0)
    assert_same(ga, oa, keys=("node_tests", "triangle_tests"))


def test_two_level_scaled_rotated_instances(device):
    m = MeshArrays()
    rng = np.random.default_rng(8)
    c = rng.uniform(-1, 1, (300, 1, 3))
    m.add_triangles(c + rng.uniform(-0.1, 0.1, (300, 3, 3)))
    mats = [trs(rng.uniform(-4, 4, 3), rng.normal(size=3), rng.uniform(0, 360), rng.uniform(0.2, 2.0))
            for _ in range(64)]
    s = m.scene(np.zeros(64, np.int32), np.stack(mats))
    s.build_accel()
    org = np.zeros((30000, 4), np.float32)
    org[:, :3] = rng.uniform(-6, 6, (30000, 3))
    org[:, 3] = 1e-5
    dd = rng.normal(size=(30000, 3))
    d = np.zeros((30000, 4), np.float32)
    d[:, :3] = dd / np.linalg.norm(dd, axis=1, keepdims=True)
    d[:, 3] = np.inf
    g = gpu_closest(device, s, org, d)
    v, i, p = s.geometry()
    tl, tp, inst = s.tlas()
    o = O.trace_two_level(v, i, p, [s.blas(0)], inst, tl, tp, org, d)
    assert_same(g, o, keys=("t", "u", "v", "triangle", "instance", "node_tests", "triangle_tests"))


@pytest.mark.parametrize("kind", ["flat", "two_level"])
def test_origin_order_gives_identical_results(device, kind):
    """Large launches claim rays in origin-cell order; every output is the
    same as with input order (bit for bit, counters included)."""
    from yaoray_amd import native as N
    if kind == "flat":
        s = Scene.triangles(20000, seed=5)
    else:
        s = Scene.instanced(instances=200, subdivisions=2, variants=4, env_width=8, env_height=4, seed=6)
    s.build_accel()
    o1, d1 = random_rays(70000, seed=9)
    o1[:, :3] = o1[:, :3] * 3.0 - 1.0  # some origins outside the scene box
    o2, d2 = s.camera_rays(256, 128)
    org, d = np.concatenate([o1, o2]), np.concatenate([d1, d2])
    org[5, 0] = np.nan  # a NaN origin falls in the first cell and misses
    device.upload_scene(s)
    try:
        device.set_trace_order(N.TRACE_ORDER_INPUT)
        a = device.trace_closest(org, d, counters=True)
        aa = device.trace_any(org, d, counters=True)
        device.set_trace_order(N.TRACE_ORDER_ORIGIN)
        b = device.trace_closest(org, d, counters=True)
        bb = device.trace_any(org, d, counters=True)
    finally:
        device.set_trace_order(N.TRACE_ORDER_ORIGIN)
    assert_same(b, a, keys=("t", "u", "v", "triangle", "instance", "node_tests", "triangle_tests"))
    for k in ("occluded", "node_tests", "triangle_tests"):
        assert np.array_equal(bb[k], aa[k]), k
    assert b["triangle"][5] == -1
    assert 0.01 < (a["triangle"] >= 0).mean() < 1.0
