"""Batched multi-camera rendering (one pass schedule for many cameras, one
HIP-graph-captured pass replayed per pass): every job's FP64 film equals the
film of rendering that job alone, bit for bit, with and without the graph,
for any wavefront size (passes may hold several sample layers of one job),
also for a textured scene with alpha cut-outs (per-job camera-ray
footprints); counters add up; every pass but the last is full."""
import numpy as np
import pytest

from yaoray_amd import Scene

pytestmark = pytest.mark.gpu


def cameras(scene, specs):
    """RenderDescs for (eye, target, fov_deg, width, height, spp) specs, with the
    host's std::tan for tan(fov/2)."""
    out = []
    for eye, target, fov, w, h, spp in specs:
        f = np.asarray(target, np.float64) - np.asarray(eye, np.float64)
        f /= np.linalg.norm(f)
        r = np.cross(f, [0.0, 1.0, 0.0])
        r /= np.linalg.norm(r)
        u = np.cross(r, f)
        scene.set_camera(eye, f, r, u, np.radians(fov))
        out.append(scene.render_desc(w, h, spp))
    return out


SPECS = [((0.0, 1.0, 3.5), (0.0, 1.0, 0.0), 40.0, 48, 32, 6),
         ((0.5, 1.2, 3.0), (0.0, 0.8, 0.0), 55.0, 20, 17, 3),
         ((-0.6, 1.5, 2.5), (0.2, 0.5, -0.5), 35.0, 37, 29, 9)]


def single(device, scene, desc, depth):
    scene.set_camera(list(desc.camera_origin), list(desc.camera_forward), list(desc.camera_right),
                     list(desc.camera_up), 0.0)
    device.film_reset(desc.width, desc.height)
    for k in range(desc.spp):
        d = scene.render_desc(desc.width, desc.height, desc.spp, depth, sample_index=k)
        d.tan_half_fov = desc.tan_half_fov
        device.render_pass(d)
    device.synchronize()
    return device.film_read(desc.width, desc.height), device.render_stats()


@pytest.mark.parametrize("kind", ["cornell", "two_level", "textured"])
def test_batch_films_equal_single_renders(device, kind):
    if kind == "cornell":
        s = Scene.cornell(1)
        depth = 8
    elif kind == "textured":
        # EWA / trilinear footprints from each job's own camera differentials,
        # alpha skips in the traversal, normal maps
        from tests.helpers import textured_scene

        s = textured_scene()
        depth = 5
    else:
        s = Scene.instanced(instances=40, subdivisions=2, variants=3, env_width=64, env_height=32, seed=5)
        depth = 10
    s.build_accel()
    device.upload_scene(s)
    device.upload_shading(s)
    jobs = cameras(s, SPECS)
    refs, totals = [], {}
    for j in jobs:
        f, st = single(device, s, j, depth)
        refs.append(f)
        for k, v in st.items():
            totals[k] = totals.get(k, 0) + v
    for graph in (True, False):
        for ppp in (0, 1000, 2500):
            films, ms, passes, st = device.render_batch(jobs, max_depth=depth, paths_per_pass=ppp, use_graph=graph)
            for k, (f, r) in enumerate(zip(films, refs)):
                for field in r.dtype.names:
                    assert np.array_equal(f[field], r[field]), (graph, ppp, k, field)
            assert st == totals, (graph, ppp)
            total = sum(j.width * j.height * j.spp for j in jobs)
            if ppp:
                assert passes == -(-total // ppp)  # every pass but the last is full
            assert np.all(ms > 0)
    # One pass holds several sample layers of a job when they fit.
    _, ms, passes_big, _ = device.render_batch(jobs, max_depth=depth, paths_per_pass=100000)
    assert passes_big == 1 and np.allclose(ms, ms[0])


def test_autotune_picks_the_fastest_probe_and_caches_it(device, tmp_path):
    from yaoray_amd import autotune

    s = Scene.cornell(1)
    s.build_accel()
    device.upload_scene(s)
    device.upload_shading(s)
    jobs = cameras(s, SPECS)
    cache = str(tmp_path / "tune.json")
    best, trials = autotune.tune(device, jobs, "cornell_3", max_depth=8, paths_candidates=(1000, 4000),
                                 blocks_candidates=(4, 8), cache=cache)
    assert len(trials) == 4
    assert best == max(trials, key=lambda r: r["paths_per_s"])
    for t in trials:
        # Each probe holds about two passes of its candidate size.
        assert t["probe_paths"] >= min(2 * t["paths_per_pass"], sum(j.width * j.height * j.spp for j in jobs))
        assert t["paths_per_s"] > 0
    assert autotune.lookup(device, "cornell_3", cache) == best
    device.set_trace_config(-1, 256)  # back to the automatic launch shape


def test_uneven_camera_schedule_against_oracle(device):
    """The batched workload's shape at reduced size: 32 cameras of uneven size
    (pixel counts log-uniform 16^2..240x135) and spp (powers of two 1..32) on a
    reduced two-level environment-lit scene, depth 12, many passes of a fixed
    wavefront (jobs interleave, several layers of a job share a pass).

    * every pixel of every job holds exactly its job's spp samples;
    * the batch's counters equal the sum over the jobs rendered one at a time
      (the schedule traces exactly each job's paths);
    * three jobs of at least 8 spp (the smallest, a middle one, the largest)
      agree with the oracle's render of the same camera, size and spp: image mean within 1 %,
      per-pixel RMSE below 0.35 x the oracle's per-pixel standard error
      (statistical parity, as for single renders)."""
    from oracle import oracle_py as O
    from yaoray_amd.api import uneven_cameras

    s = Scene.instanced(instances=300, subdivisions=2, variants=6, env_width=128, env_height=64, seed=4)
    s.build_accel()
    device.upload_scene(s)
    device.upload_shading(s)
    depth = 12
    cams = uneven_cameras(32, seed=5, min_pixels=16 * 16, max_pixels=240 * 135, min_log2_spp=0, max_log2_spp=5)
    jobs = []
    for w, h, spp, cam in cams:
        s.set_camera(*cam)
        jobs.append(s.render_desc(w, h, spp, depth))
    total = sum(j.width * j.height * j.spp for j in jobs)
    ppp = 1 << 16
    films, ms, passes, st = device.render_batch(jobs, max_depth=depth, paths_per_pass=ppp)
    print(f"uneven schedule: {len(jobs)} jobs, {total} samples, {passes} passes, {st['rays_traced']} rays", flush=True)
    assert passes == -(-total // ppp)
    for j, f in zip(jobs, films):
        assert np.all(f["samples"] == j.spp)
    totals = {}
    for j in jobs:
        _, one = device.render_batch([j], max_depth=depth, paths_per_pass=ppp)[::3]
        for k, v in one.items():
            totals[k] = totals.get(k, 0) + v
    assert st == totals
    # Jobs with at least 8 samples per pixel (a per-pixel standard error needs a few).
    order = sorted((k for k in range(len(jobs)) if jobs[k].spp >= 8),
                   key=lambda k: jobs[k].width * jobs[k].height * jobs[k].spp)
    lum = np.array([0.2126, 0.7152, 0.0722])
    for k in (order[0], order[len(order) // 2], order[-1]):
        w, h, spp, cam = cams[k]
        s.set_camera(*cam)
        o, _ = O.render(s, w, h, spp, max_depth=depth)
        g = films[k]
        gl = (g["sum"] / g["samples"][:, None]) @ lum
        ol = (o["sum"] / np.maximum(o["samples"], 1)[:, None]) @ lum
        var = o["luminance_m2"] / np.maximum(o["samples"] - 1, 1)
        noise = float(np.sqrt(np.mean(np.maximum(var, 0) / np.maximum(o["samples"], 1))))
        rmse = float(np.sqrt(np.mean((gl - ol) ** 2)))
        rel = abs(gl.mean() - ol.mean()) / ol.mean()
        print(f"job {k} {w}x{h}x{spp}: rel mean {rel:.5f} rmse {rmse:.5f} noise {noise:.5f}", flush=True)
        assert rel < 0.01 and rmse < 0.35 * noise
