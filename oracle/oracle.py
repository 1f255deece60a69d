"""ctypes bindings of the CPU oracle (liboracle.so, built from vxo_passes.cpp).

TEST INFRASTRUCTURE ONLY.  The oracle is the checker of the HIP passes: only
tests/, __graft_entry__.smoke() and the CPU-baseline leg of bench.py load it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REPO = HERE.parent
LIB_PATH = HERE / "liboracle.so"
SOURCES = ["vxo_passes.cpp"]
FLAGS = ["-std=c++17", "-O2", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-shared", "-Wall", f"-I{REPO}"]


def build(verbose: bool = False) -> Path:
    """Compiles the oracle with g++ (IEEE binary32, no contraction)."""
    tmp = LIB_PATH.with_suffix(".so.tmp")
    cmd = ["g++", *FLAGS, "-o", str(tmp), *[str(HERE / s) for s in SOURCES], "-lpthread"]
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


class Scene(C.Structure):
    _fields_ = [("chunks", C.c_void_p), ("n_chunks", C.c_uint64), ("nodes", C.c_void_p), ("ropes", C.c_void_p),
                ("n_nodes", C.c_uint64), ("grid", C.c_void_p), ("n_grid", C.c_uint64), ("macro_words", C.c_void_p),
                ("n_macro", C.c_uint64), ("colinfo", C.c_void_p), ("n_colinfo", C.c_uint64), ("clip", C.c_void_p),
                ("clip_height", C.c_void_p), ("n_clip", C.c_uint64)]


PROFILE_NAMES = ("voxel", "grass", "hdr", "fog", "shadow")


class Counts(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("grid", "meta", "macro", "node", "rope", "colinfo", "clip", "tex_rd",
                                          "tex_wr", "oob")] + [("prof", C.c_uint64 * 5)]

    def as_dict(self) -> dict:
        return {n: int(getattr(self, n)) for n, _ in self._fields_ if n != "prof"}

    def profile(self) -> dict:
        """The primary pass's profile counters (cam.profile_flags & 1), as u32 like the reference's."""
        return {n: int(v) & 0xFFFFFFFF for n, v in zip(PROFILE_NAMES, self.prof)}


def _load() -> C.CDLL:
    if not LIB_PATH.exists():
        build()
    lib = C.CDLL(str(LIB_PATH))
    vp, u32, f32 = C.c_void_p, C.c_uint32, C.c_float
    lib.vxo_primary.argtypes = [C.POINTER(Scene), vp, u32, u32, vp, vp, vp, vp, vp, vp, vp, vp, u32, u32, C.c_int,
                                C.POINTER(Counts)]
    lib.vxo_local_taa.argtypes = [u32, u32, vp, vp, vp]
    lib.vxo_query_rays.argtypes = [C.POINTER(Scene), vp, vp, u32, vp, C.c_int, C.POINTER(Counts)]
    lib.vxo_godray.argtypes = [C.POINTER(Scene), vp, u32, u32, C.c_int, vp, vp, vp, C.c_int, C.POINTER(Counts)]
    lib.vxo_composite.argtypes = [vp, u32, u32, u32, u32, vp, vp, vp, vp, vp, C.c_int]
    lib.vxo_composite_taa.argtypes = [vp, u32, u32, u32, u32, vp, vp, vp, vp, vp]
    for n in ("vxo_sin", "vxo_cos", "vxo_exp", "vxo_log"):
        getattr(lib, n).argtypes = [f32]
        getattr(lib, n).restype = f32
    lib.vxo_pow.argtypes = [f32, f32]
    lib.vxo_pow.restype = f32
    lib.vxo_hash_u32.argtypes = [u32]
    lib.vxo_hash_u32.restype = u32
    for n in ("vxo_hash12", "vxo_value_noise", "vxo_fbm"):
        getattr(lib, n).argtypes = [f32, f32]
        getattr(lib, n).restype = f32
    lib.vxo_tonemap.argtypes = [vp, vp]
    lib.vxo_phase_mie.argtypes = [f32]
    lib.vxo_phase_mie.restype = f32
    lib.vxo_intersect_aabb.argtypes = [vp, vp, vp, vp, vp]
    lib.vxo_fog_od.argtypes = [f32, vp, vp, f32]
    lib.vxo_fog_od.restype = f32
    lib.vxo_tile_raw_counts.argtypes = [C.POINTER(Scene), vp, u32, u32, vp]
    return lib


LIB = _load()


def _p(a):
    return None if a is None else np.ascontiguousarray(a).ctypes.data


class World:
    """Keeps numpy copies of the resident-world arrays alive for the oracle."""

    def __init__(self, nodes, ropes, meta, grid, macro, colinfo, clip, clip_height):
        self.arrays = [np.ascontiguousarray(a) for a in (nodes, ropes, meta, grid, macro, colinfo, clip_height)]
        self.clip = clip
        n, r, m, g, mc, ci, ch = self.arrays
        self.scene = Scene(m.ctypes.data, len(m), n.ctypes.data, r.ctypes.data, len(n), g.ctypes.data, len(g),
                           mc.ctypes.data, len(mc), ci.ctypes.data, len(ci), C.addressof(clip), ch.ctypes.data, ch.size)

    @classmethod
    def from_host(cls, hw, clip, clip_height):
        return cls(hw.nodes, hw.ropes, hw.meta, hw.grid, hw.macro, hw.colinfo, clip, clip_height)


class Frame:
    """Oracle textures of one view, with the reference's ping-pong histories."""

    def __init__(self, rw, rh, pw=None, ph=None):
        pw, ph = pw or rw, ph or rh
        self.rw, self.rh, self.pw, self.ph = rw, rh, pw, ph
        z4 = lambda h, w: np.zeros((h, w, 4), np.float32)
        self.primary_hist = z4(rh, rw)
        self.shadow = np.zeros((rh, rw), np.float32)
        self.local_hist = z4(rh, rw)
        self.godray_hist = z4(rh, rw)
        self.out_hist = z4(ph, pw)
        self.color = z4(rh, rw)
        self.depth = np.zeros((rh, rw), np.float32)
        self.local = z4(rh, rw)
        self.godray = z4(rh, rw)
        self.pre_taa = z4(ph, pw)
        self.output = z4(ph, pw)
        self.hits = None


def primary(world: World, cam, fr: Frame, hits_dtype=None, tiles=(0, 0), threads: int = 0, counts: Counts | None = None):
    """main_primary; updates fr.primary_hist / fr.shadow in place (ping-pong)."""
    hist_out = np.zeros_like(fr.primary_hist)
    shadow_out = np.zeros_like(fr.shadow)
    hits = np.zeros((fr.rh, fr.rw), hits_dtype) if hits_dtype is not None else None
    rc = LIB.vxo_primary(C.byref(world.scene), C.addressof(cam), fr.rw, fr.rh, _p(fr.primary_hist), _p(fr.shadow),
                         _p(fr.color), _p(fr.depth), _p(fr.local), _p(hist_out), _p(shadow_out),
                         None if hits is None else hits.ctypes.data, tiles[0], tiles[1], threads,
                         C.byref(counts) if counts is not None else None)
    assert rc == 0
    fr.primary_hist, fr.shadow, fr.hits = hist_out, shadow_out, hits


def query_rays(world: World, cam, rays: np.ndarray, hits_dtype, threads: int = 0, counts: Counts | None = None):
    """First voxel-geometry hit of each ray; rays is an (n, 8) float32 array (ox, oy, oz, tmax, dx, dy, dz, 0)."""
    rays = np.ascontiguousarray(rays, np.float32)
    hits = np.zeros(len(rays), hits_dtype)
    assert LIB.vxo_query_rays(C.byref(world.scene), C.addressof(cam), rays.ctypes.data, len(rays), hits.ctypes.data,
                              threads, C.byref(counts) if counts is not None else None) == 0
    return hits


def tile_raw_counts(world: World, cam, w: int, h: int) -> np.ndarray:
    """Per 8x8 tile (row-major), the raw candidate-chunk count of the tile
    gather (before the 192 cap and the 24-entry truncation)."""
    out = np.zeros(((h + 7) // 8, (w + 7) // 8), np.uint32)
    assert LIB.vxo_tile_raw_counts(C.byref(world.scene), C.addressof(cam), w, h, out.ctypes.data) == 0
    return out


def local_taa(fr: Frame):
    out = np.zeros_like(fr.local_hist)
    assert LIB.vxo_local_taa(fr.rw, fr.rh, _p(fr.local), _p(fr.local_hist), _p(out)) == 0
    fr.local_hist = out


def godray(world: World, cam, fr: Frame, macro_compat: bool = True, threads: int = 0, counts: Counts | None = None):
    out = np.zeros_like(fr.godray_hist)
    assert LIB.vxo_godray(C.byref(world.scene), C.addressof(cam), fr.rw, fr.rh, 1 if macro_compat else 0, _p(fr.depth),
                          _p(fr.godray_hist), _p(out), threads, C.byref(counts) if counts is not None else None) == 0
    fr.godray = out
    fr.godray_hist = out


def composite(cam, fr: Frame, threads: int = 0):
    assert LIB.vxo_composite(C.addressof(cam), fr.rw, fr.rh, fr.pw, fr.ph, _p(fr.color), _p(fr.godray), _p(fr.depth),
                             _p(fr.local_hist), _p(fr.pre_taa), threads) == 0


def composite_taa(cam, fr: Frame):
    out_hist = np.zeros_like(fr.out_hist)
    assert LIB.vxo_composite_taa(C.addressof(cam), fr.rw, fr.rh, fr.pw, fr.ph, _p(fr.pre_taa), _p(fr.depth),
                                 _p(fr.out_hist), _p(fr.output), _p(out_hist)) == 0
    fr.out_hist = out_hist


def frame(world: World, cam, fr: Frame, hits_dtype=None, threads: int = 0, macro_compat: bool = True):
    """One full frame in reference pass order (render/state/mod.rs:329-446)."""
    primary(world, cam, fr, hits_dtype, threads=threads)
    local_taa(fr)
    godray(world, cam, fr, macro_compat, threads=threads)
    composite(cam, fr, threads=threads)
    composite_taa(cam, fr)
