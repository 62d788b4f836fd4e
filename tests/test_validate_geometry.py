"""Device-free geometry checks (yrh_validate_geometry, run by every
yrh_upload_geometry): the traversal stack is sized from each BVH's declared
max_depth, so a declaration below the tree's real depth (root = 1, as the
reference's builder counts it, src/accel/bvh_builder.cpp:226-244) is rejected
instead of silently dropping stack entries."""
import ctypes as C

import numpy as np

from yaoray_amd import Scene
from yaoray_amd import native as N
from yaoray_amd.api import validate_geometry


def real_depth(nodes):
    depth = np.zeros(len(nodes), np.int64)
    depth[0] = 1
    for i, n in enumerate(nodes):
        if n["primitive_count"] == 0:
            for c in (i + 1, n["payload_offset"]):
                depth[c] = max(depth[c], depth[i] + 1)
    return int(depth.max())


def test_flat_bvh_depth_is_checked():
    s = Scene.triangles(3000, seed=2)
    s.build_accel()
    g = s.geometry_desc()
    assert validate_geometry(g) == (0, "")
    bvh = s.flat_bvh()
    deepest = real_depth(bvh.nodes)
    assert bvh.max_depth == deepest  # the builder's declaration is exact
    g.flat.max_depth = deepest - 1
    rc, msg = validate_geometry(g)
    assert rc == N.E_INVALID_ARGUMENT and "max_depth" in msg and str(deepest) in msg
    g.flat.max_depth = deepest + 3  # over-declaring only costs stack space
    assert validate_geometry(g)[0] == 0


def test_two_level_depths_are_checked():
    s = Scene.instanced(instances=40, subdivisions=2, variants=3, env_width=8, env_height=4, seed=5)
    s.build_accel()
    g = s.geometry_desc()
    assert validate_geometry(g) == (0, "")
    blases = (N.BvhDesc * g.blas_count).from_buffer_copy(C.string_at(g.blases, g.blas_count * C.sizeof(N.BvhDesc)))
    g.blases = C.cast(blases, C.c_void_p)
    assert validate_geometry(g)[0] == 0
    blases[1].max_depth -= 1
    rc, msg = validate_geometry(g)
    assert rc == N.E_INVALID_ARGUMENT and "max_depth" in msg
    blases[1].max_depth += 1
    g.tlas.max_depth = 1
    assert validate_geometry(g)[0] == N.E_INVALID_ARGUMENT


def test_malformed_trees_are_rejected():
    s = Scene.triangles(500, seed=3)
    s.build_accel()
    g = s.geometry_desc()
    nodes = s.flat_bvh().nodes.copy()
    interior = int(np.flatnonzero(nodes["primitive_count"] == 0)[0])
    nodes[interior]["payload_offset"] = interior  # right child not after its parent
    g.flat.nodes = nodes.ctypes.data
    rc, msg = validate_geometry(g)
    assert rc == N.E_INVALID_ARGUMENT and "child" in msg
    g = s.geometry_desc()
    g.sphere_count = 1  # a count with no sphere array
    rc, msg = validate_geometry(g)
    assert rc == N.E_INVALID_ARGUMENT and "sphere" in msg
