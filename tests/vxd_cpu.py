"""The product's device pass code built for the host (tests/cpu_kernels):
builds libvxd_cpu.so with g++ on first use and runs the primary pass in its
one-kernel and load-counting forms with host loops in place of the GPU's
tiles and lanes."""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REPO = HERE.parent
SRC = HERE / "cpu_kernels" / "vxd_cpu.cpp"
OUT = HERE / "cpu_kernels" / "_build" / "libvxd_cpu.so"
CSRC = REPO / "voxel-engine-v0_amd" / "csrc"
DEPS = [SRC, CSRC / "vxd_passes.inc", CSRC / "vxd_math.h", CSRC / "vxd_consts.h", CSRC / "vxd_scene.h",
        REPO / "include" / "voxhip.h", HERE / "cpu_kernels" / "shim" / "hip" / "hip_runtime.h"]
MODES = {"one": 0, "counted": 2, "one_k2": 3}


def build() -> Path:
    if OUT.exists() and all(OUT.stat().st_mtime >= d.stat().st_mtime for d in DEPS):
        return OUT
    OUT.parent.mkdir(parents=True, exist_ok=True)
    tmp = OUT.with_suffix(".so.tmp")
    cmd = ["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-shared", "-w",
           f"-I{HERE / 'cpu_kernels' / 'shim'}", f"-I{REPO / 'include'}", f"-I{CSRC}", "-o", str(tmp), str(SRC),
           "-lpthread"]
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        _LIB = C.CDLL(str(build()))
        vp = C.c_void_p
        _LIB.vxd_cpu_primary.argtypes = [vp, vp, C.c_uint32, C.c_uint32, vp, vp, vp, vp, vp, vp, vp, vp, C.c_int, C.c_int,
                                         vp, vp]
        _LIB.vxd_cpu_primary.restype = C.c_int
        _LIB.vxd_cpu_tile_lists.argtypes = [vp, vp, C.c_uint32, C.c_uint32, C.c_int, vp, vp, vp, vp, vp]
        _LIB.vxd_cpu_tile_lists.restype = C.c_int
        _LIB.vxd_cpu_query.argtypes = [vp, vp, vp, C.c_uint32, vp, C.c_int]
        _LIB.vxd_cpu_query.restype = C.c_int
        _LIB.vxd_cpu_convert.argtypes = [vp, C.c_uint32, vp, vp]
        _LIB.vxd_cpu_floor_quot.argtypes = [vp, vp, C.c_uint32, vp]
        _LIB.vxd_cpu_floor_quot.restype = C.c_int
        _LIB.vxd_cpu_convert.restype = C.c_int
        _LIB.vxd_cpu_math.argtypes = [vp, vp, C.c_uint32, vp]
        _LIB.vxd_cpu_math.restype = C.c_int
        _LIB.vxd_cpu_noise.argtypes = [vp, vp, C.c_uint32, vp]
        _LIB.vxd_cpu_noise.restype = C.c_int
    return _LIB


def primary(world, cam, fr, hits_dtype, mode: str = "one", threads: int = 0, profile: np.ndarray | None = None):
    """Same contract as oracle.primary: updates fr in place; returns the load counts for mode 'counted'.
    profile (5 uint64, mode 'one'): receives the profile counters when cam.profile_flags & 1."""
    threads = threads or min(16, os.cpu_count() or 1)
    hist_out = np.zeros_like(fr.primary_hist)
    shadow_out = np.zeros_like(fr.shadow)
    hits = np.zeros((fr.rh, fr.rw), hits_dtype)
    counts = np.zeros(9 + 8, np.uint64)
    rc = lib().vxd_cpu_primary(C.addressof(world.scene), C.addressof(cam), fr.rw, fr.rh, fr.primary_hist.ctypes.data,
                               fr.shadow.ctypes.data, fr.color.ctypes.data, fr.depth.ctypes.data, fr.local.ctypes.data,
                               hist_out.ctypes.data, shadow_out.ctypes.data, hits.ctypes.data, MODES[mode], threads,
                               counts.ctypes.data, None if profile is None else profile.ctypes.data)
    assert rc == 0
    fr.primary_hist, fr.shadow, fr.hits = hist_out, shadow_out, hits
    return counts


def query_rays(world, cam, rays: np.ndarray, hits_dtype, threads: int = 0):
    """The device query code run on the host: same contract as oracle.query_rays."""
    threads = threads or min(16, os.cpu_count() or 1)
    rays = np.ascontiguousarray(rays, np.float32)
    hits = np.zeros(len(rays), hits_dtype)
    assert lib().vxd_cpu_query(C.addressof(world.scene), C.addressof(cam), rays.ctypes.data, len(rays),
                               hits.ctypes.data, threads) == 0
    return hits


def floor_quot(x: np.ndarray, d: np.ndarray) -> np.ndarray:
    """The device code's floor of x / d (floor_quot), elementwise in float32."""
    x = np.ascontiguousarray(x, np.float32)
    d = np.ascontiguousarray(np.broadcast_to(np.asarray(d, np.float32), x.shape), np.float32)
    out = np.zeros(len(x), np.float32)
    assert lib().vxd_cpu_floor_quot(x.ctypes.data, d.ctypes.data, len(x), out.ctypes.data) == 0
    return out


def convert(x: np.ndarray):
    """The device code's saturating f32 -> i32 and f32 -> u32 conversions."""
    x = np.ascontiguousarray(x, np.float32)
    oi = np.zeros(len(x), np.int32)
    ou = np.zeros(len(x), np.uint32)
    assert lib().vxd_cpu_convert(x.ctypes.data, len(x), oi.ctypes.data, ou.ctypes.data) == 0
    return oi, ou


def math(x: np.ndarray, y: np.ndarray) -> np.ndarray:
    """The device code's sin, cos, exp, log and pow(x, y), as an (n, 5) array."""
    x = np.ascontiguousarray(x, np.float32)
    y = np.ascontiguousarray(y, np.float32)
    out = np.zeros((len(x), 5), np.float32)
    assert lib().vxd_cpu_math(x.ctypes.data, y.ctypes.data, len(x), out.ctypes.data) == 0
    return out


def noise(x: np.ndarray, y: np.ndarray) -> np.ndarray:
    """The device code's value_noise and fbm at (x, y), as an (n, 2) array."""
    x = np.ascontiguousarray(x, np.float32)
    y = np.ascontiguousarray(y, np.float32)
    out = np.zeros((len(x), 2), np.float32)
    assert lib().vxd_cpu_noise(x.ctypes.data, y.ctypes.data, len(x), out.ctypes.data) == 0
    return out


def fastdiv_mismatches(k: int, lo: int = 0, hi: int = 0x80000000, threads: int = 8) -> int:
    """Inputs in the non-negative bit-pattern range [lo, hi) where the device's
    divc<k> differs from the correctly rounded division."""
    f = lib().vxd_cpu_fastdiv_mismatches
    f.restype = C.c_uint64
    f.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32]
    return int(f(k, lo, hi, threads))


def fast_divisors() -> list:
    lib().vxd_cpu_fast_divisor.restype = C.c_float
    lib().vxd_cpu_fast_divisor.argtypes = [C.c_uint32]
    lib().vxd_cpu_fast_divisor_count.restype = C.c_uint32
    return [float(lib().vxd_cpu_fast_divisor(k)) for k in range(lib().vxd_cpu_fast_divisor_count())]


def tile_lists(world, cam, w: int, h: int, wave: bool) -> dict:
    """The tile candidate lists of a w x h view: the one-lane gather (wave=False)
    or the wavefront phases of the product kernel (wave=True)."""
    nt = ((w + 7) // 8) * ((h + 7) // 8)
    counts, cached = np.zeros(nt, np.uint32), np.zeros(nt, np.uint32)
    slots, enter = np.zeros((nt, 24), np.uint32), np.zeros((nt, 24), np.float32)
    over = np.zeros(1, np.uint32)
    assert lib().vxd_cpu_tile_lists(C.addressof(world.scene), C.addressof(cam), w, h, 1 if wave else 0,
                                    counts.ctypes.data, cached.ctypes.data, slots.ctypes.data, enter.ctypes.data,
                                    over.ctypes.data) == 0
    return {"count": counts, "cached": cached, "slots": slots, "enter": enter, "overflow": int(over[0])}
