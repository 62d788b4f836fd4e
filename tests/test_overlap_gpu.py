"""Render passes with each bounce's shadow rays on a second stream (beside the
next bounce's path-segment traversal) leave the same film and counters as the
one-stream sequence: every per-path operation keeps its order, so the FP64
films are compared byte for byte."""
import numpy as np
import pytest

from yaoray_amd import Scene
from yaoray_amd import native as N

pytestmark = pytest.mark.gpu


def render_both(device, s, w, h, spp, depth, batch):
    out = []
    for overlap in (False, True):
        device.set_render_overlap(overlap)
        try:
            out.append(device.render(s, w, h, spp, max_depth=depth, upload=not out, batch=batch, seed=3))
        finally:
            device.set_render_overlap(True)
    return out


@pytest.mark.parametrize("scene", ["cornell", "instanced"])
def test_overlapped_bounces_match_sequential(device, scene):
    if scene == "cornell":
        s = Scene.cornell(1)
        w, h, spp, depth, batch = 96, 64, 16, 8, 8
    else:
        # glass instances: transparent shadow walks; sky + sun environment
        s = Scene.instanced(instances=150, subdivisions=2, variants=6, env_width=64, env_height=32, seed=4)
        w, h, spp, depth, batch = 96, 64, 8, 12, 4
    s.build_accel()
    (f0, st0), (f1, st1) = render_both(device, s, w, h, spp, depth, batch)
    assert f1.tobytes() == f0.tobytes()
    assert st1 == st0
    assert f0["samples"].sum() == w * h * spp


def test_overlapped_guided_bounces_match_sequential(device):
    s = Scene.room(furniture_triangles=20000, env_width=64, env_height=32)
    s.build_accel()
    device.guide_configure([-0.1, -0.1, -0.1], [10.1, 4.1, 10.1], log2_table_size=16, min_resolution=8,
                           max_resolution=128, learning_rate=3e-3, guide_fraction=0.5, sample_capacity=1 << 16, seed=7)
    device.upload_scene(s)
    device.upload_shading(s)
    try:
        device.guide_set_mode(N.GUIDE_RECORD)
        device.render(s, 64, 36, 2, max_depth=6, upload=False, seed=9)
        device.guide_train(4, 4096)
        device.guide_set_mode(N.GUIDE_SAMPLE)
        films = []
        for overlap in (False, True):
            device.set_render_overlap(overlap)
            films.append(device.render(s, 64, 36, 8, max_depth=6, upload=False, batch=8, seed=11)[0])
    finally:
        device.set_render_overlap(True)
        device.guide_set_mode(N.GUIDE_OFF)
    assert films[1].tobytes() == films[0].tobytes()
