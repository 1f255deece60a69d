"""Oracle traversal against brute force: a dense-grid Amanatides-Woo voxel
walk in float64 (the method of the reference's own CPU pick, app/app.rs:847-961)
must find the same first solid voxel and hit distance as the oracle's SVO
traversal (ropes, macro DDA, sparse descent) for rays whose hit is not on a
cell edge or corner (ties are counted and excluded).

Geometry: random floating blocks, camera below them looking up, so the
colinfo early skip (downward rays only) and the grass probes stay inert;
wind strength 0 keeps leaf cubes undisplaced.  Candidate chunks follow the
reference's tile rule (two probe rays per 8x8 tile, ray_main.wgsl:179-197)."""
from __future__ import annotations

import math

import numpy as np
import pytest

import worlds

VS = 0.1


def _ray_dirs(cam, w, h):
    r00 = np.array(cam.ray00[:3], np.float32)
    dx = np.array(cam.ray_dx[:3], np.float32)
    dy = np.array(cam.ray_dy[:3], np.float32)
    xs = (np.arange(w, dtype=np.float32) + np.float32(0.5))
    ys = (np.arange(h, dtype=np.float32) + np.float32(0.5))
    d = r00[None, None, :] + xs[None, :, None] * dx + ys[:, None, None] * dy  # same op order as the oracle
    inv = np.float32(1.0) / np.sqrt(np.sum(d * d, axis=-1, dtype=np.float32)).astype(np.float32)
    return (d * inv[..., None]).astype(np.float32)


def _aabb(ro, rd, bmin, bmax):
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = 1.0 / rd
        t0, t1 = (bmin - ro) * inv, (bmax - ro) * inv
    te = np.max(np.minimum(t0, t1))
    tx = np.min(np.maximum(t0, t1))
    return te, tx


def _brute(ro, rd, solid, chunks, tmax=120.0):
    """First solid voxel along the ray inside the allowed chunks: (t, voxel, tie)."""
    ro = ro.astype(np.float64)
    rd = rd.astype(np.float64)
    best = None
    for (cx, cy, cz) in chunks:
        bmin = np.array([cx, cy, cz], np.float64) * 6.4
        te, tx = _aabb(ro, rd, bmin, bmin + 6.4)
        te = max(te, 0.0)
        tx = min(tx, tmax)
        if tx < te:
            continue
        p = ro + rd * (te + 1e-9)
        v = np.floor(p / VS).astype(np.int64)
        step = np.where(rd > 0, 1, -1)
        with np.errstate(divide="ignore"):
            nb = (v + (rd > 0)) * VS
            tmaxv = np.where(rd != 0, (nb - ro) / rd, np.inf)
            tdel = np.where(rd != 0, VS / np.abs(rd), np.inf)
        t = te
        lo = np.array([cx, cy, cz]) * 64
        while t <= tx:
            lv = v - lo
            if np.all(lv >= 0) and np.all(lv < 64) and solid[(cx, cy, cz)][lv[1], lv[2], lv[0]]:
                pe = ro + rd * t
                frac = pe / VS - np.round(pe / VS)
                tie = int(np.sum(np.abs(frac) < 1e-4)) >= 2
                if best is None or t < best[0]:
                    best = (t, v.copy(), tie)
                break
            a = int(np.argmin(tmaxv))
            t = tmaxv[a]
            v[a] += step[a]
            tmaxv[a] += tdel[a]
    return best


def _candidates(cam, wgx, wgy, dims):
    """Chunks crossed by the two probe rays of a tile (ray_main.wgsl:179-197)."""
    out = set()
    r00 = np.array(cam.ray00[:3], np.float64)
    dx, dy = np.array(cam.ray_dx[:3], np.float64), np.array(cam.ray_dy[:3], np.float64)
    ro = np.array(cam.cam_pos[:3], np.float64)
    for off in (4.5, 1.5):
        d = r00 + (wgx * 8 + off) * dx + (wgy * 8 + off) * dy
        d /= np.linalg.norm(d)
        for cx in range(dims[0]):
            for cy in range(dims[1]):
                for cz in range(dims[2]):
                    bmin = np.array([cx, cy, cz], np.float64) * 6.4
                    te, tx = _aabb(ro, d, bmin - 1e-6, bmin + 6.4 + 1e-6)
                    if tx >= max(te, 0.0) and te <= 120.0:
                        out.add((cx, cy, cz))
    return out


@pytest.mark.parametrize("seed,dims", [(0, (1, 1, 1)), (1, (1, 1, 1)), (2, (2, 1, 2))])
def test_oracle_traversal_matches_brute_force(vx, seed, dims):
    from oracle import oracle as O

    rng = np.random.default_rng(seed)
    vols = {(ix, 0, iz): worlds.floating_blocks(rng) for ix in range(dims[0]) for iz in range(dims[2])}
    pw = worlds.pack(vx, dims, vols)
    clip, heights = worlds.empty_clipmap(vx)
    w, h = 96, 64
    pos = (dims[0] * 3.2 + rng.uniform(-1, 1), 0.6, dims[2] * 3.2 + rng.uniform(-1, 1))
    pose = vx.make_pose(pos, float(rng.uniform(0, 2 * math.pi)), math.radians(40.0), 70.0)
    cam = vx.camera_build(pose, (w, h), (w, h), 0, wind_strength=0.0, grid_origin=(0, 0, 0), grid_dims=dims,
                          chunk_count=pw.chunk_count)
    world = O.World(pw.nodes, pw.ropes, pw.meta, pw.grid, pw.macro, pw.colinfo, clip, heights)
    fr = O.Frame(w, h)
    cnt = O.Counts()
    O.primary(world, cam, fr, vx.HIT_DTYPE, counts=cnt)
    assert cnt.oob == 0
    solid = {k: (v.reshape(64, 64, 64) != 0) for k, v in pw.mats.items()}
    dirs = _ray_dirs(cam, w, h)
    ro = np.array(cam.cam_pos[:3], np.float32)
    checked = ties = hits = 0
    for ty in range(h // 8):
        for tx in range(w // 8):
            cand = _candidates(cam, tx, ty, dims)
            for y in range(ty * 8, ty * 8 + 8):
                for x in range(tx * 8, tx * 8 + 8):
                    b = _brute(ro, dirs[y, x], solid, sorted(cand))
                    hr = fr.hits[y, x]
                    if b is not None and b[2]:
                        ties += 1
                        continue
                    checked += 1
                    if b is None:
                        assert hr["kind"] != 1, (x, y)
                        continue
                    hits += 1
                    assert hr["kind"] == 1, (x, y, b)
                    assert abs(float(hr["t"]) - b[0]) < 2e-4, (x, y, float(hr["t"]), b[0])
                    # the hit leaf contains the brute-force voxel
                    slot, node = int(hr["slot"]), int(hr["node"])
                    key = int(pw.nodes[int(pw.meta[slot]["node_base"]) + node]["key"])
                    s = 64 >> ((key >> 18) & 7)
                    v0 = np.array(hr["voxel"])
                    assert np.all(b[1] >= v0) and np.all(b[1] < v0 + s), (x, y, b[1], v0, s)
    print(f"[traversal] seed={seed} dims={dims}: checked={checked} hits={hits} ties={ties}")
    assert hits > 200 and ties < checked // 20
