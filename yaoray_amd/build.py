"""In-tree build of the native libraries (gfx950 HIP + C++20 host).

    python -m yaoray_amd.build          # libyrh.so + libyaoray_host.so

Objects go to ``_build/`` (git-ignored), libraries to ``yaoray_amd/lib/``.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
LIB = os.path.join(PKG, "lib")
OBJ = os.path.join(ROOT, "_build")

HIP_SOURCES = ["yrh_context.hip", "yrh_trace.hip", "yrh_render.hip", "yrh_guide.hip", "yrh_order.hip",
               "yrh_noalloc.hip"]
HOST_SOURCES = ["transform.cpp", "bvh_builder.cpp", "light_tables.cpp", "scenes.cpp", "abi_marshal.cpp",
                "hip_backend.cpp", "standalone_boundary.cpp", "host_capi.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# Packed FP32 ops are disabled: hipcc would otherwise emit v_pk_* f32 forms.
# FP contraction is off: traversal arithmetic follows the reference op by op.
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-command-line-argument",
             "-fvisibility=hidden"]
HOST_FLAGS = ["-std=c++20", "-O3", "-DNDEBUG", "-fPIC", "-ffp-contract=off", "-Wall", "-Wextra",
              "-Wno-unused-parameter", "-fvisibility=hidden", "-fvisibility-inlines-hidden"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(line for line in r.stdout.splitlines() if "packed-fp32-ops" not in line)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    return out


def build_hip(jobs: int = 8) -> str:
    os.makedirs(LIB, exist_ok=True)
    os.makedirs(OBJ, exist_ok=True)
    objs = []
    cmds = []
    for src in HIP_SOURCES:
        obj = os.path.join(OBJ, src + ".o")
        objs.append(obj)
        cmds.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(PKG, "csrc", src), "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, cmds))
    out = os.path.join(LIB, "libyrh.so")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", out + ".tmp", *objs, "-L/opt/rocm/lib",
          "-lrocprofiler-sdk-roctx"])
    os.replace(out + ".tmp", out)
    return out


def build_host(jobs: int = 8) -> str:
    os.makedirs(LIB, exist_ok=True)
    os.makedirs(OBJ, exist_ok=True)
    objs = []
    cmds = []
    for src in HOST_SOURCES:
        obj = os.path.join(OBJ, src + ".o")
        objs.append(obj)
        cmds.append(["g++", *HOST_FLAGS, "-c", os.path.join(PKG, "host", src), "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, cmds))
    out = os.path.join(LIB, "libyaoray_host.so")
    _run(["g++", "-shared", "-o", out + ".tmp", *objs, f"-L{LIB}", "-lyrh", "-Wl,-rpath,$ORIGIN", "-lpthread"])
    os.replace(out + ".tmp", out)
    return out


def build_all(jobs: int = 8):
    build_hip(jobs)
    build_host(jobs)


if __name__ == "__main__":
    build_all(int(sys.argv[1]) if len(sys.argv) > 1 else 8)
    print("built", os.listdir(LIB))
