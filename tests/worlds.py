"""Small hand-made worlds for CPU and GPU tests: arbitrary material volumes
per chunk, built with the product builder and packed like the resident world
(slots in grid order, contiguous node ranges, streaming/manager/slots.rs:231-251)."""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class PackedWorld:
    nodes: np.ndarray
    ropes: np.ndarray
    meta: np.ndarray
    macro: np.ndarray
    colinfo: np.ndarray
    grid: np.ndarray
    grid_origin: tuple
    grid_dims: tuple
    chunk_count: int
    mats: dict


def pack(vx, dims, volumes: dict, grid_origin=(0, 0, 0), dirt_depth=4) -> PackedWorld:
    """volumes[(ix, iy, iz)] = uint8[64*64*64] (y-slab layout); missing chunks are air."""
    nx, ny, nz = dims
    nodes, ropes, metas, macros, colinfos, mats = [], [], [], [], [], {}
    base = 0
    grid = np.full(nx * ny * nz, 0xFFFFFFFF, np.uint32)
    slot = 0
    for iz in range(nz):
        for iy in range(ny):
            for ix in range(nx):
                mat = volumes.get((ix, iy, iz))
                if mat is None:
                    mat = np.zeros(64 ** 3, np.uint8)
                mats[(ix, iy, iz)] = mat
                ground = np.full(64 * 64, -100000, np.int32)  # disables the stone collapse
                b = vx.build_chunk(mat, ground, (grid_origin[1] + iy) * 64, dirt_depth)
                m = np.zeros(1, vx.META_DTYPE)
                m["origin"][0, :3] = [(grid_origin[0] + ix) * 64, (grid_origin[1] + iy) * 64, (grid_origin[2] + iz) * 64]
                m["node_base"], m["node_count"] = base, len(b.nodes)
                m["macro_base"], m["colinfo_base"] = slot * 16, slot * 2048
                m["macro_empty"] = 0 if b.macro.any() else 1
                nodes.append(b.nodes)
                ropes.append(b.ropes)
                metas.append(m)
                macros.append(b.macro)
                colinfos.append(b.colinfo)
                grid[(iz * ny + iy) * nx + ix] = slot
                base += len(b.nodes)
                slot += 1
    return PackedWorld(np.concatenate(nodes), np.concatenate(ropes), np.concatenate(metas), np.concatenate(macros),
                       np.concatenate(colinfos), grid, tuple(grid_origin), tuple(dims), slot, mats)


def floating_blocks(rng, n_boxes=14, y_lo=24, mats=(2, 3, 4, 5)) -> np.ndarray:
    """Random axis-aligned blocks floating in the upper part of a chunk."""
    m = np.zeros((64, 64, 64), np.uint8)
    for _ in range(n_boxes):
        x0, z0 = rng.integers(0, 60, 2)
        y0 = rng.integers(y_lo, 60)
        sx, sy, sz = rng.integers(1, 14, 3)
        m[y0:y0 + sy, z0:z0 + sz, x0:x0 + sx] = rng.choice(mats)
    speck = rng.random((64, 64, 64)) < 0.002
    speck[:y_lo] = False
    m[speck] = rng.choice(mats)
    return m.reshape(-1)


def empty_clipmap(vx):
    clip = vx.Clipmap()
    clip.levels, clip.res = 0, 256
    return clip, np.zeros((1, 256, 256), np.float32)


def boundary_rays():
    """Origins on chunk corners and voxel-lattice points (fp32 sums of the
    chunk origin and whole voxels: on or within an ulp of node faces, edges and
    corners), some inside the ground and some in open air, times directions
    that tie slabs: axis-aligned, face and body diagonals, and near-parallel
    ones whose small components sit either side of the 1e-8 parallel cut
    (common.wgsl:442-444, ray/aabb.wgsl:16-84)."""
    f = np.float32
    rng = np.random.default_rng(23)
    origins = []
    for _ in range(160):
        c = rng.integers(0, (16, 4, 16))
        v = rng.integers(0, 64, 3)
        lattice = rng.integers(0, 4)  # 0: chunk corner, 1: voxel corner, 2: node-8 corner, 3: one face off-lattice
        if lattice == 0:
            v[:] = 0
        elif lattice == 2:
            v &= ~7
        o = c.astype(f) * f(6.4) + v.astype(f) * f(0.1)
        if lattice == 3:
            o[int(rng.integers(0, 3))] += f(0.05)
        origins.append(o.astype(f))
    dirs = []
    for s in [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1),
              (1, 1, 0), (1, -1, 0), (-1, 0, -1), (0, -1, 1), (1, 1, 1), (-1, -1, 1), (1, -1, -1), (-1, -1, -1),
              (1, 5e-9, 0), (1, -5e-9, 2e-8), (3e-9, -1, -3e-9), (2e-8, -1, 0), (1, -1, 1e-9), (0.5, -1, 0.25)]:
        d = np.array(s, np.float32)
        dirs.append((d / np.linalg.norm(d)).astype(f))
    rays = np.zeros((len(origins) * len(dirs), 8), np.float32)
    k = 0
    for o in origins:
        for d in dirs:
            rays[k, 0:3] = o
            rays[k, 3] = 51.2 if k % 3 else 200.0
            rays[k, 4:7] = d
            k += 1
    return rays
