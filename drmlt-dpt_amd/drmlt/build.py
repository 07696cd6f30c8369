"""Builds libdrmlt_hip.so for gfx950 with hipcc: each translation unit is
compiled to an object in parallel, then linked in-tree next to the package."""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(PKG_ROOT, "csrc")
OBJ = os.path.join(PKG_ROOT, "_build")
LIB = os.path.join(PKG_ROOT, "libdrmlt_hip.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
FLAGS = ["-O3", "-std=c++17", "-fPIC", "-munsafe-fp-atomics", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result"]
UNITS = ["k_plain.hip", "k_loop.hip", "k_count.hip", "k_deriv.hip", "k_deriv64.hip", "drmlt_capi.hip", "bvh_build.cpp", "host_io.cpp", "scene_load.cpp",
         "serialized.cpp", "bvh_device.hip"]


def _compile(unit):
    src = os.path.join(CSRC, unit)
    obj = os.path.join(OBJ, unit + ".o")
    cmd = [HIPCC, f"--offload-arch={ARCH}", *FLAGS, "-c", src, "-o", obj]
    if unit.endswith(".cpp"):
        cmd = [HIPCC, *FLAGS, "-x", "c++", "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stderr[-6000:]}")
    return obj


def build(jobs=None):
    os.makedirs(OBJ, exist_ok=True)
    jobs = jobs or min(len(UNITS), os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, UNITS))
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", LIB + ".tmp", *objs, "-lz"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stderr[-6000:]}")
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build())
