"""Autotuner for the batched multi-camera pipeline.

Times a short version of a batch (every job at `probe_spp` samples) for each
candidate wavefront size (paths per pass) and traversal launch shape (blocks
per CU), keeps the fastest, applies its launch shape to the device and, when a
cache path is given, stores the choice per device and workload key in a JSON
file so later runs start from it.
"""
from __future__ import annotations

import ctypes as C
import json
import os
import time

from . import native as N

DEFAULT_PATHS = (1 << 20, 1 << 21, 1 << 22, 1 << 23)
DEFAULT_BLOCKS_PER_CU = (4, 6, 8)


def cache_path() -> str:
    return os.environ.get("YAORAY_AMD_AUTOTUNE",
                          os.path.join(os.path.expanduser("~"), ".cache", "yaoray_amd", "autotune.json"))


def _load(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def lookup(device, key, path=None):
    """The stored choice for (device, key), or None."""
    data = _load(path or cache_path())
    return data.get(device.info()[0], {}).get(key)


def tune(device, jobs, key, max_depth=8, sampler=N.SAMPLER_OWEN_SOBOL, probe_spp=2,
         paths_candidates=DEFAULT_PATHS, blocks_candidates=DEFAULT_BLOCKS_PER_CU, cache=None, repeats=1):
    """Returns (best, trials); best = dict(paths_per_pass, blocks_per_cu, seconds)."""
    probe = []
    for j in jobs:
        d = N.RenderDesc()
        C.memmove(C.byref(d), C.byref(j), C.sizeof(N.RenderDesc))
        d.spp = min(j.spp, probe_spp)
        probe.append(d)
    total = sum(j.width * j.height * j.spp for j in probe)
    trials = []
    for bpc in blocks_candidates:
        device.set_trace_config(bpc, 0)
        for ppp in paths_candidates:
            if ppp > max(paths_candidates[0], total):
                continue  # a wavefront larger than the whole probe gains nothing
            best_t = None
            for _ in range(max(1, repeats)):
                device.synchronize()
                t0 = time.perf_counter()
                device.render_batch(probe, max_depth=max_depth, sampler=sampler, paths_per_pass=ppp)
                t = time.perf_counter() - t0
                best_t = t if best_t is None else min(best_t, t)
            trials.append({"paths_per_pass": ppp, "blocks_per_cu": bpc, "seconds": round(best_t, 5)})
    best = min(trials, key=lambda r: r["seconds"])
    device.set_trace_config(best["blocks_per_cu"], 0)
    if cache:
        data = _load(cache)
        data.setdefault(device.info()[0], {})[key] = best
        os.makedirs(os.path.dirname(os.path.abspath(cache)), exist_ok=True)
        tmp = cache + ".tmp"
        with open(tmp, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(tmp, cache)
    return best, trials
