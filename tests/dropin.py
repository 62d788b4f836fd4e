"""Runs the drop-in build (oracle/Makefile `dropin`: the reference with the
HIP backend patched into its factory) and reads the probe's output files."""
import os
import subprocess

import numpy as np

from yaoray_amd.api import FILM32_DTYPE

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DROPIN = os.path.join(ROOT, "oracle", "_ref", "dropin")
PROBE = os.path.join(DROPIN, "dropin_probe")
CLI = os.path.join(DROPIN, "yaoray")
SCENES = os.path.join(ROOT, "tests", "golden", "pbrt")
STAT_KEYS = ("rays_traced", "shadow_rays", "occluded_shadow_rays", "triangle_tests", "sphere_tests",
             "bvh_node_tests", "hits", "misses", "bad_samples", "converged_pixels")


def built():
    return os.path.exists(PROBE) and os.path.exists(CLI)


def read_result(path):
    with open(path, "rb") as f:
        data = f.read()
    assert data[:4] == b"YRPF"
    ok, w, h, spp = np.frombuffer(data, np.int32, 4, 4)
    stats = dict(zip(STAT_KEYS, (int(x) for x in np.frombuffer(data, np.uint64, 10, 20))))
    n = int(np.frombuffer(data, np.int32, 1, 100)[0])
    err = data[104:104 + n].decode()
    off = 104 + n
    film = np.frombuffer(data, FILM32_DTYPE, (len(data) - off) // FILM32_DTYPE.itemsize, off)
    return {"ok": bool(ok), "width": int(w), "height": int(h), "spp": int(spp), "stats": stats, "error": err,
            "film": film.reshape(int(h), int(w)) if len(film) else None}


def probe(args, out, cwd):
    r = subprocess.run([PROBE, *args[:3], out, *args[3:]], cwd=cwd, capture_output=True, text=True, timeout=600)
    return r, (read_result(out) if os.path.exists(out) else None)


def render(scene, backend, tmp, *extra):
    out = os.path.join(str(tmp), f"{backend}.bin")
    return probe(["render", os.path.join(SCENES, scene), backend, *extra], out, str(tmp))


def checkpoint(scene, backend, tmp, passes):
    ckpt = os.path.join(str(tmp), "film.yrc")
    out = os.path.join(str(tmp), f"{backend}_resumed.bin")
    r = subprocess.run([PROBE, "checkpoint", os.path.join(SCENES, scene), backend, ckpt, out, str(passes)],
                       cwd=str(tmp), capture_output=True, text=True, timeout=600)
    return r, (read_result(out) if os.path.exists(out) else None), ckpt
