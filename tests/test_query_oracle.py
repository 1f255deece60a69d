"""Pins the oracle's ray query (first voxel-geometry hit through the chunk
grid, oracle/vxo_trace.h query_ray) against brute force: a float64
Amanatides-Woo walk over the dense voxel grid of the whole world (the method
of the reference's CPU pick, app/app.rs:847-961) must find the same first
solid voxel, material and distance, for rays that start inside and outside the
grid, leave it, or run out of range.  Hits on a voxel edge or corner are
counted and excluded (the two methods may legitimately pick different voxels
there)."""
from __future__ import annotations

import math

import numpy as np
import pytest

import worlds

VS = 0.1


def _dense(pw_mats, dims):
    nx, ny, nz = dims
    g = np.zeros((ny * 64, nz * 64, nx * 64), np.uint8)  # [y, z, x]
    for (ix, iy, iz), m in pw_mats.items():
        g[iy * 64:(iy + 1) * 64, iz * 64:(iz + 1) * 64, ix * 64:(ix + 1) * 64] = m.reshape(64, 64, 64)
    return g


def _brute(g, origin_vox, ro, rd, tmax):
    """First solid voxel along the ray within [0, tmax]: (t, voxel, mat, tie) or None."""
    ro = ro.astype(np.float64)
    rd = rd.astype(np.float64)
    lo = np.array(origin_vox, np.float64) * VS
    hi = lo + np.array([g.shape[2], g.shape[0], g.shape[1]], np.float64) * VS
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = 1.0 / rd
        t0, t1 = (lo - ro) * inv, (hi - ro) * inv
    te = max(float(np.max(np.minimum(t0, t1))), 0.0)
    tx = min(float(np.min(np.maximum(t0, t1))), tmax)
    if not (tx >= te):
        return None
    p = ro + rd * (te + 1e-9)
    v = np.floor(p / VS).astype(np.int64)
    v = np.clip(v, np.array(origin_vox), np.array(origin_vox) + np.array([g.shape[2], g.shape[0], g.shape[1]]) - 1)
    step = np.where(rd > 0, 1, -1)
    with np.errstate(divide="ignore"):
        nb = (v + (rd > 0)) * VS
        tm = np.where(rd != 0, (nb - ro) / rd, np.inf)
        td = np.where(rd != 0, VS / np.abs(rd), np.inf)
    t = te
    shape = np.array([g.shape[2], g.shape[0], g.shape[1]])
    while t <= tx:
        lv = v - np.array(origin_vox)
        if np.any(lv < 0) or np.any(lv >= shape):
            return None
        m = g[lv[1], lv[2], lv[0]]
        if m:
            pe = ro + rd * t
            frac = pe / VS - np.round(pe / VS)
            tie = int(np.sum(np.abs(frac) < 1e-4)) >= 2
            return t, v.copy(), int(m), tie
        a = int(np.argmin(tm))
        t = tm[a]
        v[a] += step[a]
        tm[a] += td[a]
    return None


def _check(vx, world, cam, pw_nodes, pw_meta, g, origin_vox, rays, min_hits):
    from oracle import oracle as O

    cnt = O.Counts()
    hits = O.query_rays(world, cam, rays, vx.HIT_DTYPE, counts=cnt)
    assert cnt.oob == 0
    checked = ties = nhit = 0
    for i, r in enumerate(rays):
        b = _brute(g, origin_vox, r[:3], r[4:7], float(r[3]))
        h = hits[i]
        if b is not None and b[3]:
            ties += 1
            continue
        checked += 1
        if b is None:
            assert h["kind"] == 0, (i, h)
            continue
        nhit += 1
        assert h["kind"] == 1, (i, b)
        assert abs(float(h["t"]) - b[0]) < 2e-4, (i, float(h["t"]), b[0])
        assert int(h["mat"]) == b[2], (i, int(h["mat"]), b[2])
        key = int(pw_nodes[int(pw_meta[int(h["slot"])]["node_base"]) + int(h["node"])]["key"])
        s = 64 >> ((key >> 18) & 7)
        v0 = np.array(h["voxel"])
        assert np.all(b[1] >= v0) and np.all(b[1] < v0 + s), (i, b[1], v0, s)
    print(f"[query] checked={checked} hits={nhit} ties={ties}")
    assert nhit >= min_hits and ties < max(2, checked // 20)


def _cam(vx, dims, chunk_count, origin=(0, 0, 0)):
    pose = vx.make_pose((1.0, 1.0, 1.0), 0.0, 0.0, 70.0)
    return vx.camera_build(pose, (64, 64), (64, 64), 0, wind_strength=2.0, grid_origin=origin, grid_dims=dims,
                           chunk_count=chunk_count)


def _random_rays(rng, n, lo, hi, tmax_range):
    o = rng.uniform(lo, hi, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = o
    rays[:, 3] = rng.uniform(*tmax_range, n)
    rays[:, 4:7] = d.astype(np.float32)
    rays[:, 4:7] /= np.linalg.norm(rays[:, 4:7], axis=1, keepdims=True)  # unit in fp32 as well
    return rays


@pytest.mark.parametrize("seed", [0, 1])
def test_query_floating_blocks(vx, seed):
    from oracle import oracle as O

    rng = np.random.default_rng(seed)
    dims = (2, 2, 2)
    vols = {(ix, iy, iz): worlds.floating_blocks(rng, n_boxes=40, y_lo=4) for ix in range(2) for iy in range(2) for iz in range(2)
            if (ix + iy + iz) % 2 == 0}
    pw = worlds.pack(vx, dims, vols)
    clip, heights = worlds.empty_clipmap(vx)
    world = O.World(pw.nodes, pw.ropes, pw.meta, pw.grid, pw.macro, pw.colinfo, clip, heights)
    cam = _cam(vx, dims, pw.chunk_count)
    g = _dense(pw.mats, dims)
    rays = _random_rays(rng, 800, -3.0, 15.8, (2.0, 25.0))  # some start outside the 12.8 m grid
    _check(vx, world, cam, pw.nodes, pw.meta, g, (0, 0, 0), rays, min_hits=60)


def test_query_seeded_world(vx):
    """The seeded terrain world (8x4x8 chunks): rays from above the surface."""
    from oracle import oracle as O

    from conftest import host_world

    sc, hw = host_world("c0")
    p = sc.params
    dims = sc.grid_dims
    mats = {}
    for iz in range(dims[2]):
        for iy in range(dims[1]):
            for ix in range(dims[0]):
                mats[(ix, iy, iz)] = vx.fill_chunk(p, ix, iy, iz)[0]
    g = _dense(mats, dims)
    clip, heights = worlds.empty_clipmap(vx)
    world = O.World.from_host(hw, clip, heights)
    cam = _cam(vx, dims, hw.chunk_count)
    rng = np.random.default_rng(5)
    rays = _random_rays(rng, 300, 1.0, 50.0, (3.0, 20.0))
    rays[:, 1] = rng.uniform(8.0, 16.0, len(rays))  # around the surface (y = 4.8 .. 14.4 m)
    _check(vx, world, cam, hw.nodes, hw.meta, g, (0, 0, 0), rays, min_hits=100)
