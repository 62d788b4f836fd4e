"""Autotuner for the batched multi-camera pipeline.

For each candidate wavefront size (paths per pass) and traversal launch shape
(blocks per CU) it renders a probe of about two passes of that size (every
job at the same reduced sample count) and measures paths per second: a pass
ends when its slowest paths do, so the rate depends on the pass size and a
probe much smaller than a pass would not show it. The fastest rate wins; its
launch shape is applied to the device and, when a cache path is given, the
choice is stored per device and workload key in a JSON file.
"""
from __future__ import annotations

import ctypes as C
import json
import os
import time

from . import native as N
from .api import YrhError

# Up to 2^28 paths per pass (~128 GB of path state at ~480 B per path, within
# one MI355X's 288 GB); a candidate the device cannot hold is skipped.
DEFAULT_PATHS = (1 << 23, 1 << 24, 1 << 25, 1 << 26, 1 << 27, 1 << 28)
DEFAULT_BLOCKS_PER_CU = (-1, 4)  # -1: the automatic launch shape


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


def tune(device, jobs, key, max_depth=8, sampler=N.SAMPLER_OWEN_SOBOL, paths_candidates=DEFAULT_PATHS,
         blocks_candidates=DEFAULT_BLOCKS_PER_CU, cache=None, repeats=1, passes_per_probe=2):
    """Returns (best, trials); best = dict(paths_per_pass, blocks_per_cu, paths_per_s)."""
    pixels = sum(j.width * j.height for j in jobs)
    full = sum(j.width * j.height * j.spp for j in jobs)
    trials = []
    for ppp in paths_candidates:
        if ppp > 2 * full:
            continue  # larger than the whole schedule
        spp = max(1, -(-passes_per_probe * ppp // pixels))
        probe = []
        for j in jobs:
            d = N.RenderDesc()
            C.memmove(C.byref(d), C.byref(j), C.sizeof(N.RenderDesc))
            d.spp = min(j.spp, spp)
            probe.append(d)
        paths = sum(d.width * d.height * d.spp for d in probe)
        for bpc in blocks_candidates:
            device.set_trace_config(bpc, 0)
            best_t = None
            failed = None
            for _ in range(max(1, repeats)):
                device.synchronize()
                t0 = time.perf_counter()
                try:
                    device.render_batch(probe, max_depth=max_depth, sampler=sampler, paths_per_pass=ppp)
                except YrhError as e:  # e.g. the pass's path state does not fit the device
                    failed = str(e)
                    break
                t = time.perf_counter() - t0
                best_t = t if best_t is None else min(best_t, t)
            if failed is not None:
                trials.append({"paths_per_pass": ppp, "blocks_per_cu": bpc, "probe_paths": paths, "error": failed})
                continue
            trials.append({"paths_per_pass": ppp, "blocks_per_cu": bpc, "probe_paths": paths,
                           "seconds": round(best_t, 5), "paths_per_s": round(paths / best_t, 1)})
    best = max((r for r in trials if "paths_per_s" in r), key=lambda r: r["paths_per_s"])
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
