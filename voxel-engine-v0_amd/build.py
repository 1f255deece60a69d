"""Build libvoxhip.so (HIP kernels for gfx950 + C ABI + host producers).

Usage: python voxel-engine-v0_amd/build.py [--verbose]

The library is written next to the Python bindings, in
voxel-engine-v0_amd/voxhip/libvoxhip.so, so that it travels with the tree.
Objects go to voxel-engine-v0_amd/_build/.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
REPO = PKG.parent
CSRC = PKG / "csrc"
OUT_DIR = PKG / "voxhip"
BUILD = PKG / "_build"
LIB = OUT_DIR / "libvoxhip.so"

ARCH = os.environ.get("VOXHIP_ARCH", "gfx950")

# Every HIP translation unit: IEEE fp32 with no contraction, correctly
# rounded divide/sqrt, and no packed-fp32 instructions.
HIP_FLAGS = [
    "-O3",
    "-std=c++17",
    f"--offload-arch={ARCH}",
    "-fPIC",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    f"-I{REPO / 'include'}",
    f"-I{CSRC}",
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-ffp-contract=off", f"-I{REPO / 'include'}", f"-I{CSRC}"]

HIP_SOURCES = ["vx_kernels.hip", "vx_runtime.hip", "vx_server.hip", "vx_builder.hip", "vx_sort.hip"]
CXX_SOURCES = ["host_world.cpp"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def run(cmd: list[str], verbose: bool) -> None:
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd[:3])} ... (exit {r.returncode})")
    if verbose and (r.stdout or r.stderr):
        sys.stderr.write(r.stdout + r.stderr)


def build(verbose: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    hc = hipcc()
    jobs = []
    objs = []
    for src in HIP_SOURCES:
        obj = BUILD / (src + ".o")
        jobs.append([hc, *HIP_FLAGS, "-c", str(CSRC / src), "-o", str(obj)])
        objs.append(obj)
    for src in CXX_SOURCES:
        obj = BUILD / (src + ".o")
        jobs.append(["g++", *CXX_FLAGS, "-c", str(CSRC / src), "-o", str(obj)])
        objs.append(obj)
    with cf.ThreadPoolExecutor(max_workers=min(4, len(jobs))) as ex:
        list(ex.map(lambda c: run(c, verbose), jobs))
    tmp = LIB.with_suffix(".so.tmp")
    run([hc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs), "-lpthread"], verbose)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--verbose", action="store_true")
    a = ap.parse_args()
    print(build(a.verbose))
