#!/usr/bin/env python3
"""Closest-hit ray-casting throughput on one MI355X.

Workload (the ray-cast microbench): 1,000,000 random triangles (centres
U[0,1]^3, vertices c + l*d, d uniform on the sphere, l log-uniform in
[1e-3, 2e-2], PCG32 seed 2) in the reference's own SAH BVH (built by the host
library, byte-identical to yr::BuildBvh); 2048x2048 coherent pinhole rays
(eye (0.5,0.5,-1.5) -> (0.5,0.5,0.5), 40 degree fov, pixel centres) plus
4,194,304 incoherent rays (origins U[0,1]^3, directions uniform, seed 3);
t_min 1e-5, t_max +inf. One step = one closest-hit pass over all 8,388,608
rays, rays and tree resident in HBM. Prints one JSON line on stdout.

After it, the render workloads (`renders` in the JSON line): the seeded
Cornell box (256x256, 16 spp, depth 8) and the two-level instanced scene
(1920x1080, 64 spp, depth 16); one step = one whole frame, timed with events
on the context stream, beside the reference's RenderCpuPathTrace on the host
cores (the whole Cornell frame; a 480x270 x 8 spp sample of the two-level one);
the adaptive Cornell render; and two PBRT scenes as the reference's frontend
compiles them, rendered by HipBackend: the reference's own Cornell scene and
the textured scene of the drop-in tests at 1280x720 x 64 spp.

Then (`batch`) 32 cameras of uneven size and spp on the two-level scene in
one pass schedule with HIP-graph replay, wavefront size autotuned; and
(`guided_vs_plain`) the equal-time relMSE of the guided and the plain
integrator on the window-lit room, guided time including its training.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mrays/s (closest-hit, 1 GPU) + equal-time MSE, guided vs plain wavefront"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def ensure_built():
    from yaoray_amd import native

    if not os.path.exists(native.lib_path("libyrh.so")) or not os.path.exists(native.lib_path("libyaoray_host.so")):
        from yaoray_amd import build

        build.build_all()


def algorithmic_bytes(n_rays, node_tests, tri_tests, out_bytes):
    # 32 B ray in + result out + 32 B per node box test + 36 B per triangle test.
    return 32.0 * n_rays + out_bytes * n_rays + 32.0 * float(node_tests) + 36.0 * float(tri_tests)


def load_traffic(kernel_key):
    path = os.path.join(ROOT, "profiles", "traffic_c2_microbench.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        data = json.load(f)
    entry = data.get(kernel_key)
    if not entry:
        return None
    # rocprofv3 FETCH_SIZE / WRITE_SIZE (KiB per dispatch), each scaled by what
    # the counter reads for a device copy of known size.
    return round(entry["fetch_kib"] * 1024.0 * data["fetch_factor"] + entry["write_kib"] * 1024.0 *
                 data["write_factor"])


def load_counters():
    """Cache, VALU and occupancy figures of the closest-hit kernel from the
    committed rocprofv3 counter summary (profiles/counters_c2_microbench.json)."""
    path = os.path.join(ROOT, "profiles", "counters_c2_microbench.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        data = json.load(f)
    keys = ("l2_hit_rate", "l1_hit_rate", "valu_busy", "valu_lane_utilization", "occupancy",
            "valu_instructions_per_wave", "salu_instructions_per_wave")
    out = {k: data[k] for k in keys if k in data}
    if "valu_busy" in out and "valu_lane_utilization" in out:
        # Issue-bound figure: the fraction of the CU's vector lane-cycles doing work.
        out["valu_lane_cycles_used"] = round(out["valu_busy"] * out["valu_lane_utilization"], 4)
    if out.get("valu_instructions_per_wave"):
        out["salu_per_valu"] = round(out.get("salu_instructions_per_wave", 0) / out["valu_instructions_per_wave"], 3)
    return out


def host_cores():
    """(threads used, nproc, cores in this process's affinity). The CPU baselines
    use every core of the affinity mask up to 16: one GPU's share of an
    8-GPU MI355X host's cores, whatever nproc reports."""
    aff = len(os.sched_getaffinity(0))
    return min(16, aff), os.cpu_count(), aff


def median_seconds(fn, repeats=5):
    """One untimed warm-up call, then the median wall time of `repeats` calls
    (the reference's benchmark convention, README.md:90); returns (median s,
    all timings, last result)."""
    fn()
    times, out = [], None
    for _ in range(repeats):
        t0 = time.perf_counter()
        out = fn()
        times.append(time.perf_counter() - t0)
    return float(np.median(times)), [round(t, 4) for t in times], out


def timed_passes(dev, launch, steps):
    """Per-launch kernel milliseconds from events on the context stream, plus wall seconds."""
    starts = [dev.event() for _ in range(steps)]
    stops = [dev.event() for _ in range(steps)]
    dev.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        dev.record(starts[k])
        launch()
        dev.record(stops[k])
    dev.synchronize()
    wall = time.perf_counter() - t0
    ms = [dev.elapsed_ms(a, b) for a, b in zip(starts, stops)]
    for e in starts + stops:
        dev.destroy_event(e)
    return ms, wall


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--triangles", type=int, default=1_000_000)
    ap.add_argument("--resolution", type=int, default=2048)
    ap.add_argument("--incoherent", type=int, default=4_194_304)
    ap.add_argument("--cpu-stride", type=int, default=4, help="CPU baseline traces every k-th ray")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--renders", default="cornell,two_level,cornell_adaptive,cornell_pbrt,textures_pbrt",
                    help="render workloads timed after the ray-cast microbench (comma list, empty for none)")
    ap.add_argument("--render-frames", type=int, default=3)
    ap.add_argument("--no-batch", action="store_true", help="skip the batched multi-camera workload")
    ap.add_argument("--no-batch-graph", action="store_true",
                    help="batched passes as plain stream launches instead of HIP-graph replay (same kernels, films)")
    ap.add_argument("--no-guiding", action="store_true", help="skip the guided-vs-plain comparison")
    ap.add_argument("--guide-width", type=int, default=1280)
    ap.add_argument("--guide-height", type=int, default=720)
    ap.add_argument("--guide-ref-spp", type=int, default=16384)
    ap.add_argument("--guide-spp", default="16,32,64,128,256,512")
    ap.add_argument("--guide-train-passes", type=int, default=8)
    ap.add_argument("--guide-record-spp", type=int, default=2, help="sample indices per recording pass")
    ap.add_argument("--guide-depths", type=int, default=2, help="guided bounces (0 = all that may continue)")
    ap.add_argument("--guide-fraction", type=float, default=0.65)
    ap.add_argument("--guide-train-steps", type=int, default=16)
    ap.add_argument("--guide-train-batch", type=int, default=16384)
    ap.add_argument("--guide-lr", type=float, default=3e-3)
    ap.add_argument("--guide-threshold", type=float, default=0.1,
                    help="relMSE threshold of the headline time-to-threshold figure (0.2 is reported beside it)")
    ap.add_argument("--blocks-per-cu", type=int, default=0)
    ap.add_argument("--lds-levels", type=int, default=0, help="traversal stack levels kept in LDS (1..16)")
    ap.add_argument("--block-size", type=int, default=0)
    ap.add_argument("--trace-order", choices=("origin", "input"), default="origin",
                    help="order in which the ray-cast kernels claim rays (origin cells, or input order)")
    ap.add_argument("--calibration-copy-mb", type=int, default=0,
                    help="also run a device copy of this size (HBM counter calibration under rocprofv3)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (run with --gpus 1)",
              file=sys.stderr)
        sys.exit(2)

    ensure_built()
    from yaoray_amd import Device, Scene, random_rays

    t = time.perf_counter()
    scene = Scene.triangles(args.triangles, seed=2)
    build_s = scene.build_accel()
    info = scene.info
    log(f"scene {args.triangles} tris, {info.node_count} nodes, depth {info.max_depth}, build {build_s:.2f}s")
    o1, d1 = scene.camera_rays(args.resolution, args.resolution)
    o2, d2 = random_rays(args.incoherent, seed=3)
    org = np.ascontiguousarray(np.concatenate([o1, o2]))
    dirs = np.ascontiguousarray(np.concatenate([d1, d2]))
    n = len(org)
    n_coherent = len(o1)
    log(f"{n} rays ready in {time.perf_counter() - t:.1f}s")

    dev = Device(0)
    name, cus = dev.info()
    if args.blocks_per_cu or args.block_size:
        dev.set_trace_config(args.blocks_per_cu, args.block_size)
    if args.lds_levels:
        dev.set_trace_lds_levels(args.lds_levels)
    from yaoray_amd import native as N
    dev.set_trace_order(N.TRACE_ORDER_ORIGIN if args.trace_order == "origin" else N.TRACE_ORDER_INPUT)
    dev.upload_scene(scene)
    d_org, d_dir = dev.to_device(org), dev.to_device(dirs)
    d_hit, d_inst = dev.alloc(16 * n), dev.alloc(4 * n)
    d_nt, d_tt, d_occ = dev.alloc(4 * n), dev.alloc(4 * n), dev.alloc(4 * n)

    # Counting passes: the per-ray node/triangle test counts (equal to the
    # reference's counters; checked against it on the CPU sample below).
    dev.trace_closest_device(d_org, d_dir, n, d_hit, d_inst, d_nt, d_tt)
    dev.synchronize()
    nt_closest = d_nt.download(np.uint32, (n,)).astype(np.int64)
    tt_closest = d_tt.download(np.uint32, (n,)).astype(np.int64)
    hits = d_hit.download(np.float32, (n, 4))
    dev.trace_any_device(d_org, d_dir, n, d_occ, d_nt, d_tt)
    dev.synchronize()
    nt_any = d_nt.download(np.uint32, (n,)).astype(np.int64).sum()
    tt_any = d_tt.download(np.uint32, (n,)).astype(np.int64).sum()
    occluded = d_occ.download(np.uint32, (n,))
    hit_ids = hits[:, 3].view(np.int32)
    any_consistent = bool(np.array_equal(occluded.astype(bool), hit_ids >= 0))
    log(f"counting pass: hit rate {(hit_ids >= 0).mean():.4f}, mean node tests {nt_closest.mean():.1f}, "
        f"triangle tests {tt_closest.mean():.1f}; any==closest.hit {any_consistent}")

    def closest():
        dev.trace_closest_device(d_org, d_dir, n, d_hit)

    def anyhit():
        dev.trace_any_device(d_org, d_dir, n, d_occ)

    for _ in range(args.warmup):
        closest()
    dev.synchronize()
    log(f"warmup {args.warmup} done; timing {args.steps} steps")
    ms_closest, wall = timed_passes(dev, closest, args.steps)
    for _ in range(max(1, args.warmup)):
        anyhit()
    ms_any, wall_any = timed_passes(dev, anyhit, args.steps)
    if args.calibration_copy_mb > 0:
        # A plain elementwise copy of known size: under rocprofv3 --pmc its
        # FETCH_SIZE / WRITE_SIZE give the counter-to-bytes factor.
        import torch

        src = torch.ones((args.calibration_copy_mb << 20) // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"calibration copy of {args.calibration_copy_mb} MiB done")

    # Phase split of the same call (ray-ordering pass, traversal kernel), from
    # events the library records around each on the context stream.
    dev.set_trace_timing(True)
    phases = []
    for _ in range(args.steps):
        closest()
        phases.append(dev.trace_phase_ms())
    dev.set_trace_timing(False)
    order_ms = float(np.mean([p[0] for p in phases]))
    trace_ms = float(np.mean([p[1] for p in phases]))

    value = n * args.steps / wall / 1e6
    avg_closest_s = float(np.mean(ms_closest)) / 1e3
    avg_any_s = float(np.mean(ms_any)) / 1e3
    b_closest = algorithmic_bytes(n, nt_closest.sum(), tt_closest.sum(), 16)
    b_any = algorithmic_bytes(n, nt_any, tt_any, 4)
    achieved = b_closest / (trace_ms / 1e3) / 1e9
    log(f"closest: {value:.1f} Mrays/s wall, call {avg_closest_s * 1e3:.3f} ms (ordering {order_ms:.3f} ms, "
        f"traversal kernel {trace_ms:.3f} ms, {achieved:.0f} GB/s algorithmic); any: {n / avg_any_s / 1e6:.1f} Mrays/s")

    cpu = None
    parity = None
    if not args.no_cpu_baseline:
        cpu, parity = cpu_baseline(scene, org, dirs, hits, nt_closest, tt_closest, args.cpu_stride, n_coherent)

    renders = {}
    for key in [k for k in args.renders.split(",") if k]:
        if key in PBRT_WORKLOADS:
            renders[key] = bench_pbrt(key, args.render_frames, host_cores()[0], not args.no_cpu_baseline)
            continue
        renders[key] = bench_render(dev, key, args.render_frames, min(16, len(os.sched_getaffinity(0))),
                                    not args.no_cpu_baseline)
        if key == "two_level":
            renders[key]["binding"] = two_level_binding()
    batch = (bench_batch(dev, not args.no_cpu_baseline, host_cores()[0], use_graph=not args.no_batch_graph)
             if not args.no_batch else None)
    guiding = None
    if not args.no_guiding:
        guiding = bench_guiding(dev, args.guide_width, args.guide_height, args.guide_ref_spp,
                                [int(v) for v in args.guide_spp.split(",")], args.guide_train_passes,
                                args.guide_train_steps, args.guide_train_batch, args.guide_lr, args.guide_threshold,
                                args.guide_record_spp, args.guide_depths, args.guide_fraction)

    traffic = load_traffic("closest")
    result = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "ray-cast microbench: 1M random triangles (reference SAH BVH), 2048x2048 coherent pinhole "
                        "+ 4194304 incoherent rays, closest-hit, t in (1e-5, inf]",
            "triangles": args.triangles,
            "rays": n,
            "coherent_rays": n_coherent,
            "incoherent_rays": n - n_coherent,
            "bvh_nodes": info.node_count,
            "bvh_max_depth": info.max_depth,
            "device": name,
            "compute_units": cus,
            "parallelism": "single GPU, persistent lanes with work-counter refill",
            "ray_order": args.trace_order,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "kernel": "TraceFlatKernel<false> (closest hit)",
            "algorithmic_bytes_per_launch": b_closest,
            "avg_kernel_ms": round(trace_ms, 4),
            "avg_call_ms": round(avg_closest_s * 1e3, 4),
            "avg_ordering_ms": round(order_ms, 4),
            "bytes_model": "32 B ray + 16 B hit + 32 B per node_test + 36 B per triangle_test",
            "mean_node_tests": round(float(nt_closest.mean()), 3),
            "mean_triangle_tests": round(float(tt_closest.mean()), 3),
            "hbm_achieved": round(traffic / (trace_ms / 1e3) / 1e9, 1) if traffic else None,
            "hbm_frac": round(traffic / (trace_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4) if traffic else None,
            "counters": load_counters(),
            "binding": "VALU issue: valu_lane_cycles_used = valu_busy x lane utilization is the share of vector "
                       "lane-cycles doing work; HBM is not the bound (hbm_frac), nor L2 (hit rate)",
            "note": "achieved prices every node and triangle visit as an HBM fetch (32 B / 36 B); L2 and the "
                    "Infinity Cache serve most revisits, so achieved can exceed the HBM peak; traffic is the "
                    "measured HBM bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, calibrated), hbm_frac "
                    "that traffic over the kernel time against the 8 TB/s peak; counters from the committed "
                    "rocprofv3 summary of the same kernel",
        },
        "cpu_baseline": cpu,
        "any_hit": {
            "value": round(n / avg_any_s / 1e6, 2),
            "unit": "Mrays/s (kernel time)",
            "wall_value": round(n * args.steps / wall_any / 1e6, 2),
            "avg_kernel_ms": round(avg_any_s * 1e3, 4),
            "algorithmic_GBps": round(b_any / avg_any_s / 1e9, 1),
            "occluded_equals_closest_hit": any_consistent,
        },
        "closest_kernel_mrays": round(n / avg_closest_s / 1e6, 2),
        "parity_sample": parity,
        "renders": renders,
        "batch": batch,
        "guided_vs_plain": guiding,
    }
    print(json.dumps(result), flush=True)


RENDER_WORKLOADS = {
    "cornell": dict(desc="seeded Cornell box (34 tris, ceiling emitter L=17), 256x256, 16 spp OwenSobol, depth 8",
                    width=256, height=256, spp=16, max_depth=8,
                    cpu=dict(width=256, height=256, spp=16)),
    "two_level": dict(desc="ground + 10k instanced 4x-subdivided icospheres (51.2M instanced tris; diffuse / GGX "
                           "conductor / dielectric), 2048x1024 sky+sun env, 1920x1080, 64 spp OwenSobol, depth 16",
                      width=1920, height=1080, spp=64, max_depth=16,
                      cpu=dict(width=480, height=270, spp=8)),
    # Adaptive sampling with the reference's default settings
    # (AdaptiveSamplingSettings: min_spp 16, relative 0.05, absolute 0.001, confidence 1.96).
    "cornell_adaptive": dict(desc="seeded Cornell box, 256x256, up to 256 spp OwenSobol, depth 8, adaptive sampling "
                                  "(min 16 spp, relative error 0.05, absolute 0.001, confidence 1.96)",
                             width=256, height=256, spp=256, max_depth=8, adaptive=(16, 0.05, 0.001, 1.96),
                             cpu=dict(width=256, height=256, spp=256)),
}


# Samples per active pixel in one adaptive pass (ray budget vs pass count: a
# one-sample pass over the few live pixels of a late pass is latency bound).
ADAPTIVE_BATCH = 16


def batch_jobs(scene, count=32, seed=5, cameras=None):
    """Cameras of the batched workload (yaoray_amd.api.uneven_cameras): pixel counts
    log-uniform in [128^2, 1920x1080] at 16:9 (clamped), spp log-uniform over powers
    of two in {4..256}, eyes on a ring 25..50 units out looking at points near the
    centre, fov 30..60 degrees; depth 16."""
    from yaoray_amd.api import uneven_cameras

    jobs = []
    for w, h, spp, cam in uneven_cameras(count, seed):
        scene.set_camera(*cam)
        jobs.append(scene.render_desc(w, h, spp, 16))
        if cameras is not None:
            cameras.append(cam)
    return jobs


def two_level_binding():
    """What bounds the two-level frame (and the batched pipeline, which runs the
    same kernels): the committed rocprofv3 counters of its two traversal
    kernels (profiles/counters_two_level_c3.json)."""
    path = os.path.join(ROOT, "profiles", "counters_two_level_c3.json")
    try:
        with open(path) as f:
            data = json.load(f)
    except (OSError, ValueError):
        return None
    keys = ("valu_busy", "valu_lane_utilization", "occupancy", "l2_hit_rate")
    out = {"source": "profiles/counters_two_level_c3.json"}
    for k, v in data.items():
        if isinstance(v, dict) and "valu_busy" in v:
            out[k] = {x: v[x] for x in keys if x in v}
            out[k]["valu_lane_cycles_used"] = round(v["valu_busy"] * v["valu_lane_utilization"], 3)
    out["binding"] = ("VALU issue with a quarter of the lanes active (TLAS / instance entry / BLAS node / triangle "
                      "paths diverge within a wave) and launch tails (occupancy below the register and LDS limits); "
                      "not HBM (L2 hit rate 0.93-0.97)")
    return out


def bench_batch(dev, with_cpu=True, cpu_threads=16, cpu_scale=8, use_graph=True):
    """32 cameras on the two-level scene in one pass schedule with HIP-graph replay;
    wavefront size and launch shape from the autotuner. CPU baseline: the 32 jobs
    rendered one after another by the reference's RenderCpuPathTrace, each at
    1/cpu_scale of its width and height and 1/cpu_scale of its spp."""
    from yaoray_amd import Scene, autotune

    scene = Scene.instanced()
    scene.build_accel()
    cameras = []
    jobs = batch_jobs(scene, cameras=cameras)
    dev.upload_scene(scene)
    dev.upload_shading(scene)
    t = time.perf_counter()
    best, trials = autotune.tune(dev, jobs, "two_level_32_cameras", max_depth=16)
    tune_s = time.perf_counter() - t
    log(f"batch: autotune {tune_s:.1f}s -> {best}")
    films, ms, passes, st = dev.render_batch(jobs, max_depth=16, paths_per_pass=best["paths_per_pass"],
                                             use_graph=use_graph)
    dev.set_trace_config(-1, 0)  # later sections use the automatic launch shape
    total_ms = float(ms.max())
    samples = int(sum(j.width * j.height * j.spp for j in jobs))
    out = {
        "workload": "32 cameras on the two-level scene (pixels log-uniform 128^2..1920x1080 at 16:9, spp 4..256, "
                    "seed 5), depth 16, one pass schedule, "
                    + ("HIP graph replayed per pass" if use_graph else "passes launched on the stream (no graph)"),
        "mean_camera_done_ms": round(float(np.mean(ms)), 2),
        "value": round(st["rays_traced"] / (total_ms / 1e3) / 1e6, 2),
        "unit": "Mrays/s (path segments, rays_traced)",
        "shadow_Mrays_s": round(st["shadow_rays"] / (total_ms / 1e3) / 1e6, 2),
        "ms_total": round(total_ms, 2),
        "samples": samples,
        "samples_per_s": round(samples / (total_ms / 1e3), 1),
        "passes": passes,
        "paths_per_pass": best["paths_per_pass"],
        "blocks_per_cu": best["blocks_per_cu"],
        "per_camera": [{"width": j.width, "height": j.height, "spp": j.spp, "done_ms": round(float(m), 2)}
                       for j, m in zip(jobs, ms)],
        "autotune_trials": trials,
        "binding": two_level_binding(),
    }
    log(f"batch: {total_ms:.0f} ms, {out['value']} Mrays/s over {passes} passes")
    if with_cpu:
        from oracle import oracle_py as O

        _, nproc, aff = host_cores()
        rays = samples_cpu = 0
        secs = 0.0
        kind = "reference" if O.ref_available() else "port"
        for j, cam in zip(jobs, cameras):
            scene.set_camera(*cam)
            w, h, spp = max(1, j.width // cpu_scale), max(1, j.height // cpu_scale), max(1, j.spp // cpu_scale)
            if kind == "reference":
                _, rst, sec = O.ref_render(scene, w, h, spp, max_depth=16, threads=cpu_threads)
            else:
                t0 = time.perf_counter()
                _, rst = O.render(scene, w, h, spp, max_depth=16, threads=cpu_threads)
                sec = time.perf_counter() - t0
            rays += rst["rays_traced"]
            samples_cpu += w * h * spp
            secs += sec
        out["cpu_baseline"] = {
            "value": round(rays / secs / 1e6, 4),
            "unit": "Mrays/s (path segments, rays_traced)",
            "cores": cpu_threads,
            "nproc": nproc,
            "affinity_cores": aff,
            "kind": kind,
            "sample": f"the 32 jobs rendered one after another by RenderCpuPathTrace, each at 1/{cpu_scale} of its "
                      f"width and height and 1/{cpu_scale} of its spp ({samples_cpu} samples, "
                      f"{samples_cpu / samples:.4g} of the batch), render seconds summed ({secs:.2f}s)",
            "batch_seconds_extrapolated": round(secs * samples / samples_cpu, 1),
        }
        log(f"batch: cpu {kind} {out['cpu_baseline']['value']} Mrays/s")
    return out


def relmse(film, ref):
    """mean over pixels and channels of (x - r)^2 / (r^2 + 0.01)."""
    x = film["sum"] / np.maximum(film["samples"], 1)[:, None]
    r = ref["sum"] / np.maximum(ref["samples"], 1)[:, None]
    return float(np.mean((x - r) ** 2 / (r ** 2 + 0.01)))


def relmse_luminance_corrected(film, ref):
    """Luminance relMSE with the reference's own noise taken out: with x and r
    independent, E[(x - r)^2] = E[(x - mu)^2] + Var(r), and Var(r) of each pixel's
    mean is its Welford M2 / (n - 1) / n from the FP64 film."""
    lum = np.array([0.2126, 0.7152, 0.0722])
    x = (film["sum"] / np.maximum(film["samples"], 1)[:, None]) @ lum
    r = (ref["sum"] / np.maximum(ref["samples"], 1)[:, None]) @ lum
    n = ref["samples"].astype(np.float64)
    var_r = ref["luminance_m2"] / np.maximum(n - 1, 1) / np.maximum(n, 1)
    return float(np.mean(((x - r) ** 2 - var_r) / (r ** 2 + 0.01)))


def interp_loglog(xs, ys, x):
    """log-log interpolation (linear extrapolation past the ends) of y(x)."""
    lx, ly = np.log(np.asarray(xs, float)), np.log(np.asarray(ys, float))
    if x <= xs[0]:
        k = 0
    elif x >= xs[-1]:
        k = len(xs) - 2
    else:
        k = int(np.searchsorted(xs, x)) - 1
    t = (np.log(x) - lx[k]) / (lx[k + 1] - lx[k])
    return float(np.exp(ly[k] + t * (ly[k + 1] - ly[k])))


def bench_guiding(dev, width, height, ref_spp, spp_list, train_passes, train_steps, train_batch, lr, threshold,
                  record_spp=1, guide_depths=0, guide_fraction=0.5):
    """Equal-time relMSE of the guided and the plain wavefront integrator on the room
    workload against a plain FP64 reference. Each mode is one progressive render whose
    film is read at the spp checkpoints (readback excluded from the clock); the guided
    render's first `train_passes` passes (`record_spp` sample indices each) record
    training samples and each is followed by `train_steps` Adam steps, all inside its
    clock, and those passes stay in its film (every pass is an unbiased estimate)."""
    from yaoray_amd import Scene
    from yaoray_amd import native as N

    scene = Scene.room()
    scene.build_accel()
    dev.upload_scene(scene)
    dev.upload_shading(scene)
    depth = 8
    total = max(spp_list)
    npix = width * height
    big = max(1, min(total, dev.suggest_paths_per_pass(depth) // npix))

    ref, _ = dev.render(scene, width, height, 1, max_depth=depth, upload=False)  # warm-up
    dev.synchronize()
    t0 = time.perf_counter()
    ref, _ = dev.render(scene, width, height, ref_spp, max_depth=depth, upload=False, batch=big, seed=0x5EED)
    dev.synchronize()
    ref_s = time.perf_counter() - t0
    log(f"guiding: reference {ref_spp} spp in {ref_s:.1f}s")

    train_events = []

    def progressive(guided):
        dev.film_reset(width, height)
        points = []
        clock = 0.0
        done = 0
        if guided:
            dev.guide_configure([-0.1, -0.1, -0.1], [10.1, 4.1, 10.1], log2_table_size=18, min_resolution=8,
                                max_resolution=512, learning_rate=lr, guide_fraction=guide_fraction,
                                sample_capacity=1 << 22, seed=7)
            dev.guide_set_depth_limit(guide_depths)
        try:
            for checkpoint in sorted(spp_list):
                dev.synchronize()
                t = time.perf_counter()
                while done < checkpoint:
                    if guided and done < train_passes * record_spp:
                        dev.guide_set_mode(N.GUIDE_RECORD)
                        count = min(record_spp, checkpoint - done)
                    else:
                        if guided:
                            dev.guide_set_mode(N.GUIDE_SAMPLE)
                        count = min(big, checkpoint - done)
                    d = scene.render_desc(width, height, total, depth, sample_index=done)
                    d.sample_count = count
                    dev.render_pass(d)
                    done += count
                    if guided and done <= train_passes * record_spp:
                        e0, e1 = dev.event(), dev.event()
                        dev.record(e0)
                        dev.guide_train(train_steps, train_batch)
                        dev.record(e1)
                        train_events.append((e0, e1))
                dev.synchronize()
                clock += time.perf_counter() - t
                film = dev.film_read(width, height)
                points.append({"spp": checkpoint, "seconds": round(clock, 4), "relmse": relmse(film, ref),
                               "relmse_lum_corrected": relmse_luminance_corrected(film, ref)})
        finally:
            if guided:
                dev.guide_set_mode(N.GUIDE_OFF)
        return points

    plain = progressive(False)
    guided = progressive(True)
    dev.synchronize()
    train_ms = [dev.elapsed_ms(a, b) for a, b in train_events]
    for a, b in train_events:
        dev.destroy_event(a)
        dev.destroy_event(b)
    # Training's MFMA work per sample: forward (2 x (32x64 + 64x64 + 64x64 + 64x48)
    # multiply-adds) plus the backward's input and weight gradients (twice that).
    flop_per_sample = 3 * 2 * (32 * 64 + 2 * 64 * 64 + 64 * 48)
    steps_done = len(train_ms) * train_steps
    train_s = sum(train_ms) / 1e3
    training_rate = None
    if steps_done and train_s > 0:
        tflops = steps_done * train_batch * flop_per_sample / train_s / 1e12
        training_rate = {
            "kernels": "per Adam step: GuideBackwardKernel (encode, forward, mixture NLL, backward, grid-gradient "
                       "atomics), GuideWeightGradKernel, GuideAdamMlpKernel, GuideAdamGridKernel, GuidePackKernel",
            "flop_per_sample": flop_per_sample,
            "steps": steps_done,
            "batch": train_batch,
            "ms_per_step": round(sum(train_ms) / steps_done, 4),
            "achieved": round(tflops, 3),
            "peak": 2500.0,
            "unit": "TFLOP/s (BF16 MFMA, FP32 accumulation)",
            "frac": round(tflops / 2500.0, 5),
            "bound": "not MFMA: the hash-grid gradient's atomic adds and the latency of one wave per 16 samples "
                     "(2 waves per SIMD); the matrix products are a few percent of the step",
            "timing": "HIP events on the context stream around each yrh_guide_train call, inside the guided clock",
        }
    px = [p["seconds"] for p in plain]
    py = [p["relmse"] for p in plain]
    for g in guided:
        g["plain_relmse_at_equal_time"] = interp_loglog(px, py, g["seconds"])
        g["improvement"] = round(g["plain_relmse_at_equal_time"] / g["relmse"], 3)
        p = next(q for q in plain if q["spp"] == g["spp"])
        g["per_sample_improvement"] = round(p["relmse"] / g["relmse"], 3)

    def time_to(points, thr, key="relmse"):
        """Time at which the measured curve reaches thr (log-log interpolation
        between measured points only; None when thr is outside them)."""
        xs = [p["seconds"] for p in points]
        ys = [max(p[key], 1e-12) for p in points]
        for k in range(len(xs) - 1):
            if ys[k] >= thr >= ys[k + 1]:
                t = (np.log(thr) - np.log(ys[k])) / (np.log(ys[k + 1]) - np.log(ys[k]))
                return round(float(np.exp(np.log(xs[k]) + t * (np.log(xs[k + 1]) - np.log(xs[k])))), 4)
        return None

    out = {
        "workload": f"window-lit room with a hidden ceiling lamp (200k furniture tris, diffuse + GGX), "
                    f"{width}x{height}, depth 8, OwenSobol",
        "reference": f"plain, {ref_spp} spp, FP64 film, {ref_s:.1f}s",
        "relmse": "mean over pixels and channels of (x - r)^2 / (r^2 + 0.01)",
        "guide_fraction": guide_fraction,
        "guided_bounces": guide_depths if guide_depths > 0 else "all that may continue",
        "training": {"passes": train_passes, "spp_per_pass": record_spp, "steps_per_pass": train_steps,
                     "batch": train_batch,
                     "learning_rate": lr, "note": "inside the guided clock; its passes stay in the film"},
        "plain": plain,
        "guided": guided,
        "equal_time_improvement": guided[-1]["improvement"],
        "training_mfma": training_rate,
        "query_note": "GuideQueryKernel (hash-grid gathers + the forward MLP on MFMA, 26.6 kFLOP per row) runs "
                      "inside each guided bounce; its share of the frame is in profiles/guide_kernels_room.md",
        "threshold": threshold,
        "time_to_threshold_plain_s": time_to(plain, threshold),
        "time_to_threshold_guided_s": time_to(guided, threshold),
        "time_to_threshold_by_level": {str(t): {"plain_s": time_to(plain, t), "guided_s": time_to(guided, t)}
                                       for t in sorted({0.2, threshold, 0.05}, reverse=True)},
        "relmse_lum_corrected": "luminance relMSE minus the reference's own per-pixel variance of the mean "
                                "(M2 / (n - 1) / n of its FP64 film) over (r^2 + 0.01): unbiased for the error "
                                "against the converged image",
        "time_to_threshold_lum_corrected": {
            str(t): {"plain_s": time_to(plain, t, "relmse_lum_corrected"),
                     "guided_s": time_to(guided, t, "relmse_lum_corrected")}
            for t in sorted({0.2, threshold, 0.05}, reverse=True)},
    }
    tp, tg = out["time_to_threshold_plain_s"], out["time_to_threshold_guided_s"]
    out["time_to_threshold_speedup"] = round(tp / tg, 3) if tp and tg else None
    log(f"guiding: equal-time relMSE improvement {out['equal_time_improvement']}x at {guided[-1]['spp']} spp; "
        f"time to relMSE <= {threshold}: plain {tp}s, guided {tg}s")
    return out


PBRT_WORKLOADS = {
    # key: (fixture, section prefix, frame override (w, h, spp) or None, description)
    "cornell_pbrt": ("cornell_pbrt", "full_", None,
                     "the reference's Cornell scene (cornell_box_pbrt.pbrt: 24 tris, mirror + glass spheres, area + "
                     "point light)"),
    "textures_pbrt": ("textures_pbrt", "ours_", (1280, 720, 64),
                      "hip_textures.pbrt (EWA / trilinear / nearest / bilinear image textures, normal map with "
                      "tangents, textured thin glass, 16x8 HDR environment map, area light)"),
}


def bench_pbrt(key, frames, cpu_threads, with_cpu):
    """A PBRT scene as the reference's frontend compiles it (the committed
    RenderSceneIR dump under tests/golden/), rendered by HipBackend (Prepare +
    Render, the reference's pass loop); Render seconds per frame. cornell_pbrt:
    the reference's own Cornell scene (scenes/pbrt/cornell_box_pbrt: mirror and
    glass spheres, an area light and a point light; 256x256, 64 spp, depth 6,
    independent sampler); textures_pbrt: this repository's textured scene at
    1280x720, 64 spp, depth 6 (camera-ray footprints, EWA and trilinear
    lookups, normal mapping, textured shadows)."""
    from yaoray_amd import Scene

    fixture, prefix, frame, desc = PBRT_WORKLOADS[key]
    f = np.load(os.path.join(ROOT, "tests", "golden", fixture + ".npz"), allow_pickle=False)
    sections = {k[len(prefix):]: f[k] for k in f.files if k.startswith(prefix)}
    scene, st = Scene.from_ir(sections)
    w, h, spp, depth = st["width"], st["height"], st["spp"], st["max_depth"]
    if frame is not None:
        w, h, spp = frame
    # No progress callback (cancel_after_passes < 0): Render batches its passes.
    kw = dict(max_depth=depth, sampler=st["sampler"], seed=st["seed"], cancel_after_passes=-1)
    scene.backend_render(w, h, spp, **kw)  # warm-up
    reps = [scene.backend_render(w, h, spp, **kw) for _ in range(frames)]
    secs = [r[0]["elapsed_seconds"] for r in reps]
    i = int(np.argsort(secs)[len(secs) // 2])
    rep, film = reps[i]
    sec = secs[i]
    sampler = {0: "independent", 1: "stratified", 2: "owen-sobol", 3: "zsobol"}.get(st["sampler"], "?")
    out = {
        "workload": f"{desc}, {w}x{h}, {spp} spp {sampler}, depth {depth}, HipBackend Render",
        "frames": frames,
        "ms_per_frame": round(sec * 1e3, 3),
        "value": round(rep["rays_traced"] / sec / 1e6, 2),
        "unit": "Mrays/s (path segments, rays_traced)",
        "shadow_Mrays_s": round(rep["shadow_rays"] / sec / 1e6, 2),
        "rays_traced": rep["rays_traced"],
        "bad_samples": rep["bad_samples"],
        "prepare_ms": round(rep["prepare_seconds"] * 1e3, 3),
        "cpu_baseline": None,
    }
    log(f"render {key}: {out['ms_per_frame']} ms/frame, {out['value']} Mrays/s (+{out['shadow_Mrays_s']} shadow)")
    if with_cpu:
        from oracle import oracle_py as O

        if not O.ref_available():
            return out
        c_spp = max(1, spp // (4 if frame is None else 16))
        rsecs = []

        def run():
            fr, st_, sec_ = O.ref_render(scene, w, h, c_spp, max_depth=depth, sampler=st["sampler"], seed=st["seed"],
                                         threads=cpu_threads)
            rsecs.append(sec_)
            return st_

        run()
        rsecs.clear()
        rst = [run() for _ in range(5)][-1]
        rsec = float(np.median(rsecs))
        _, nproc, aff = host_cores()
        out["cpu_baseline"] = {
            "value": round(rst["rays_traced"] / rsec / 1e6, 4),
            "unit": "Mrays/s (path segments, rays_traced)",
            "cores": cpu_threads,
            "nproc": nproc,
            "affinity_cores": aff,
            "kind": "reference",
            "sample": f"{w}x{h} at {c_spp} spp of the same scene and settings, RenderCpuPathTrace render seconds, "
                      f"warm-up + median of 5 ({rsec:.3f}s)",
            "runs_s": [round(x, 4) for x in rsecs],
            "frame_seconds_extrapolated": round(rsec * spp / c_spp, 3),
        }
        log(f"render {key}: cpu reference {out['cpu_baseline']['value']} Mrays/s on {cpu_threads} threads")
    return out


def bench_render(dev, key, frames, cpu_threads, with_cpu):
    """One step = one whole frame (all spp) of the render workload; rays counted as the
    reference counts rays_traced (camera + continuation segments), shadow rays apart."""
    from yaoray_amd import Scene

    w = RENDER_WORKLOADS[key]
    t = time.perf_counter()
    scene = Scene.cornell(1) if key.startswith("cornell") else Scene.instanced()
    scene.build_accel()
    log(f"render {key}: scene ready in {time.perf_counter() - t:.2f}s")
    dev.upload_scene(scene)
    dev.upload_shading(scene)
    # Sample indices per pass: as many as the suggested paths per pass hold
    # (the whole frame fits on one MI355X for both workloads).
    batch = int(max(1, min(w["spp"], dev.suggest_paths_per_pass(w["max_depth"]) // (w["width"] * w["height"]))))
    kw = dict(max_depth=w["max_depth"], upload=False, batch=batch)
    adaptive = w.get("adaptive")
    dev.render(scene, w["width"], w["height"], batch, **kw)  # warm-up, allocates the pass buffers
    dev.synchronize()
    secs = []
    st = None
    film = None
    passes = converged = None
    for _ in range(frames):
        e0, e1 = dev.event(), dev.event()
        dev.record(e0)
        if adaptive:
            # One pass per sample index over the active pixels (the host reads the
            # active count after each pass, as the reference's loop does).
            film, st, passes, converged = dev.render_adaptive(scene, w["width"], w["height"], w["spp"], adaptive,
                                                              max_depth=w["max_depth"], upload=False,
                                                              batch=ADAPTIVE_BATCH)
        else:
            film, st = dev.render(scene, w["width"], w["height"], w["spp"], **kw)
        dev.record(e1)
        dev.synchronize()
        secs.append(dev.elapsed_ms(e0, e1) / 1e3)
        dev.destroy_event(e0)
        dev.destroy_event(e1)
    sec = float(np.median(secs))
    out = {
        "workload": w["desc"],
        "frames": frames,
        "ms_per_frame": round(sec * 1e3, 3),
        "value": round(st["rays_traced"] / sec / 1e6, 2),
        "unit": "Mrays/s (path segments, rays_traced)",
        "shadow_Mrays_s": round(st["shadow_rays"] / sec / 1e6, 2),
        "samples_per_s": round(float(film["samples"].sum()) / sec, 1),
        "rays_traced": st["rays_traced"],
        "shadow_rays": st["shadow_rays"],
        "mean_node_tests_per_ray": round(st["bvh_node_tests"] / max(1, st["rays_traced"] + st["shadow_rays"]), 2),
        "samples_per_pass": batch,
        "cpu_baseline": None,
    }
    if adaptive:
        out["samples_per_pass"] = (f"{ADAPTIVE_BATCH} per active pixel (first pass min_spp), added in order until "
                                   f"the pixel converges; rays_traced includes the samples not added")
        out["passes"] = passes
        out["converged_pixels"] = converged
        out["mean_spp"] = round(float(film["samples"].mean()), 2)
    log(f"render {key}: {out['ms_per_frame']} ms/frame, {out['value']} Mrays/s (+{out['shadow_Mrays_s']} shadow)")
    if with_cpu:
        from oracle import oracle_py as O

        c = w["cpu"]
        _, nproc, aff = host_cores()
        if O.ref_available():
            kind = "reference"
            rsecs = []

            def run():
                f, st_, sec_ = O.ref_render(scene, c["width"], c["height"], c["spp"], max_depth=w["max_depth"],
                                            threads=cpu_threads, adaptive=adaptive)
                rsecs.append(sec_)  # RenderCpuPathTrace's own render seconds (prepare excluded)
                return f, st_

            run()
            rsecs.clear()
            results = [run() for _ in range(5)]
            rsec = float(np.median(rsecs))
            rfilm, rst = results[-1]
            runs = [round(x, 4) for x in rsecs]
        else:
            kind = "port"
            if adaptive:
                rsec, runs, res = median_seconds(lambda: O.render_adaptive(
                    scene, c["width"], c["height"], c["spp"], adaptive, max_depth=w["max_depth"], threads=cpu_threads))
                rfilm, rst = res[0], res[1]
            else:
                rsec, runs, (rfilm, rst) = median_seconds(lambda: O.render(
                    scene, c["width"], c["height"], c["spp"], max_depth=w["max_depth"], threads=cpu_threads))
        frac = (c["width"] * c["height"] * c["spp"]) / (w["width"] * w["height"] * w["spp"])
        out["cpu_baseline"] = {
            "value": round(rst["rays_traced"] / rsec / 1e6, 4),
            "unit": "Mrays/s (path segments, rays_traced)",
            "cores": cpu_threads,
            "nproc": nproc,
            "affinity_cores": aff,
            "kind": kind,
            "sample": f"{c['width']}x{c['height']} at {c['spp']} spp of the same scene and settings "
                      f"({frac:.4g} of the frame's samples), RenderCpuPathTrace render seconds, warm-up + median "
                      f"of 5 ({rsec:.3f}s)",
            "runs_s": runs,
            "frame_seconds_extrapolated": round(rsec / frac, 2),
        }
        if (c["width"], c["height"], c["spp"]) == (w["width"], w["height"], w["spp"]):
            g = film["sum"] / np.maximum(film["samples"], 1)[:, None]
            r = rfilm["sum"] / np.maximum(rfilm["samples"], 1)[:, None]
            lum = np.array([0.2126, 0.7152, 0.0722])
            gl, rl = g @ lum, r @ lum
            var = rfilm["luminance_m2"] / np.maximum(rfilm["samples"] - 1, 1)
            noise = float(np.sqrt(np.mean(np.maximum(var, 0) / np.maximum(rfilm["samples"], 1))))
            rmse = float(np.sqrt(np.mean((gl - rl) ** 2)))
            out["parity"] = {
                "checked_against": kind,
                "rel_mean_diff": float(abs(g.mean() - r.mean()) / max(abs(r.mean()), 1e-30)),
                "per_pixel_rmse_luminance": rmse,
                "cpu_film_noise_luminance": noise,
                "rmse_over_noise": round(rmse / max(noise, 1e-30), 4),
                "pixels_bit_equal_fp32": float(np.mean(np.all(g.astype(np.float32) == r.astype(np.float32), axis=1))),
                "rays_traced_equal": bool(rst["rays_traced"] == st["rays_traced"]),
                "rays_traced_rel_diff": round(abs(rst["rays_traced"] - st["rays_traced"]) / max(1, rst["rays_traced"]),
                                              6),
                "note": "same sampler streams; paths coincide until a transcendental's last bit differs, so the "
                        "per-pixel RMSE is a fraction of the noise (statistical parity bound: 0.35)",
            }
            if adaptive:
                out["parity"]["pixels_with_equal_sample_counts"] = float(np.mean(film["samples"] == rfilm["samples"]))
                out["parity"]["rays_traced_vs_reference"] = round(st["rays_traced"] / max(1, rst["rays_traced"]), 4)
        log(f"render {key}: cpu {kind} {out['cpu_baseline']['value']} Mrays/s on {cpu_threads} threads")
    return out


def cpu_baseline(scene, org, dirs, hits, nt, tt, stride, n_coherent):
    """The reference's own IntersectAcceleration (oracle/_ref) on every stride-th ray,
    timed on the host cores; falls back to the restated oracle where _ref is absent."""
    from oracle import oracle_py as O

    threads, nproc, aff = host_cores()
    sel = np.arange(0, len(org), stride)
    o, d = np.ascontiguousarray(org[sel]), np.ascontiguousarray(dirs[sel])
    v, i, p = scene.geometry()
    t = time.perf_counter()
    if O.ref_available():
        kind = "reference"
        ref = O.RefScene(v, i, p)
        log(f"cpu baseline: reference BVH rebuilt in {time.perf_counter() - t:.2f}s; tracing {len(sel)} rays "
            f"on {threads} threads")
        secs, runs, r = median_seconds(lambda: ref.trace(o, d, threads=threads))
    else:
        kind = "port"
        bvh = scene.flat_bvh()
        secs, runs, r = median_seconds(lambda: O.trace_flat(v, i, p, bvh, o, d, threads=threads))
    g_id = hits[sel, 3].view(np.int32)
    g_t = hits[sel, 0]
    id_match = float(np.mean(g_id == r["triangle"]))
    both = (g_id >= 0) & (g_id == r["triangle"])
    rel = np.abs(g_t[both] - r["t"][both]) / np.maximum(np.abs(r["t"][both]), 1e-30)
    parity = {
        "rays": int(len(sel)),
        "checked_against": kind,
        "hit_id_match": id_match,
        "t_max_rel_err": float(rel.max()) if len(rel) else 0.0,
        "t_bit_exact": bool(np.array_equal(g_t.view(np.uint32), r["t"].view(np.uint32))),
        "node_tests_equal": bool(np.array_equal(nt[sel], r["node_tests"].astype(np.int64))),
        "triangle_tests_equal": bool(np.array_equal(tt[sel], r["triangle_tests"].astype(np.int64))),
    }
    n_coh = int(np.sum(sel < n_coherent))
    cpu = {
        "value": round(len(sel) / secs / 1e6, 4),
        "unit": "Mrays/s",
        "cores": threads,
        "nproc": nproc,
        "affinity_cores": aff,
        "kind": kind,
        "sample": f"every {stride}th ray of the same workload ({n_coh} coherent + {len(sel) - n_coh} incoherent), "
                  f"yr::IntersectAcceleration per ray; warm-up + median of 5 runs ({secs:.2f}s)",
        "runs_s": runs,
    }
    log(f"cpu baseline ({kind}, {threads} threads): {cpu['value']} Mrays/s; parity {parity}")
    return cpu, parity


if __name__ == "__main__":
    main()
