"""Incremental clipmap update (src/clipmap.rs:182-577): the product's planner
(vx_clipmap_update + vx_clipmap_fill_rects) against the oracle's restatement
(oracle/clipmap.py), and the oracle itself pinned by what the construction
fixes: a first touch equals a full build, a walk along one axis keeps the
texture equal to a full rebuild, a two-axis walk displaces strips by exactly
the other axis' offset (the reference's placement), a jump of res cells or more
rebuilds the level, and the cadence gate / edge guard decide when a frame
updates.  CPU only."""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import clipmap as OC

RES = 64
LEVELS = 3


@pytest.fixture(scope="module")
def world(vx):
    return vx.world_params(42)


def height_fn(p):
    return lambda vx_, vz_: OC.ground_height_np(p.seed, p.base_y, p.amplitude, p.wavelength0, vx_, vz_)


def full_logical(vx, p, cam_x, cam_z, levels=LEVELS, res=RES):
    """Per level, a from-scratch build at the camera, in logical order."""
    clip, h = vx.clipmap_build(p, cam_x, cam_z, levels=levels, res=res)
    return clip, [vx.clip_logical(clip, h, l) for l in range(levels)]


def product_step(vx, p, st, heights, cam_x, cam_z, t):
    clip, rects = st.update(cam_x, cam_z, t)
    data = vx.clipmap_fill_rects(p, rects)
    vx.apply_clip_rects(heights, rects, data)
    return clip, rects, data


def walk(seed, n=60):
    """A seeded camera walk: small steps both ways, medium strides, and two
    teleports (beyond level 0's and level 1's windows)."""
    rng = np.random.default_rng(seed)
    x, z = 211.3, -97.6
    out = []
    for k in range(n):
        if k in (20, 41):
            x += 150.0 if k == 20 else -300.0
            z += 40.0
        elif k % 9 == 4:
            x += float(rng.uniform(-30, 30))
            z += float(rng.uniform(-30, 30))
        else:
            x += float(rng.uniform(-2.5, 2.5))
            z += float(rng.uniform(-2.5, 2.5))
        out.append((x, z, 0.1 * k))
    return out


def test_ground_height_restatement(vx, world):
    rng = np.random.default_rng(5)
    xs = rng.integers(-200000, 200000, 400)
    zs = rng.integers(-200000, 200000, 400)
    got = OC.ground_height_np(world.seed, world.base_y, world.amplitude, world.wavelength0, xs, zs)
    want = [vx.ground_height(world, int(a), int(b)) for a, b in zip(xs, zs)]
    assert np.array_equal(got, np.array(want, np.int32))


@pytest.mark.parametrize("wrap_fix", [False, True])
@pytest.mark.parametrize("min_interval", [0.0, 0.25])
def test_update_matches_oracle(vx, world, wrap_fix, min_interval):
    st = vx.ClipmapState(levels=LEVELS, res=RES, min_update_interval_s=min_interval, wrap_fix=wrap_fix)
    orc = OC.Clipmap(LEVELS, RES, 1.0, min_interval, wrap_fix)
    hf = height_fn(world)
    for k, (x, z, t) in enumerate(walk(11)):
        if k == 33:
            st.invalidate()
            orc.invalidate()
        clip, rects = st.update(x, z, t)
        params, ups = orc.update(hf, x, z, t)
        data = vx.clipmap_fill_rects(world, rects)
        assert len(rects) == len(ups), (k, [r.as_tuple() for r in rects], [(u["x"], u["w"]) for u in ups])
        for r, d, u in zip(rects, data, ups):
            assert (r.level, r.x, r.y, r.w, r.h, r.lx0, r.lz0) == (u["level"], u["x"], u["y"], u["w"], u["h"], u["lx0"],
                                                                   u["lz0"])
            assert np.array_equal(d, u["data"]), (k, r.as_tuple())
        assert clip.levels == LEVELS and clip.res == RES and clip.base_cell_m == 1.0
        for l in range(LEVELS):
            assert tuple(clip.level[l][:3]) == tuple(float(v) for v in params["level"][l]), (k, l)
            assert tuple(clip.offset[l][:2]) == params["offset"][l], (k, l)
        for l in range(LEVELS, 16):
            assert tuple(clip.level[l]) == (0.0, 0.0, 1.0, 0.0) and tuple(clip.offset[l]) == (0, 0, 0, 0)


def test_first_touch_is_a_full_build(vx, world):
    orc = OC.Clipmap(LEVELS, RES)
    cx, cz = -1234.56, 789.01
    params, ups = orc.update(height_fn(world), cx, cz, 0.0)
    clip, logical = full_logical(vx, world, cx, cz)
    assert [(u["level"], u["x"], u["y"], u["w"], u["h"]) for u in ups] == [(l, 0, 0, RES, RES) for l in range(LEVELS)]
    for l, u in enumerate(ups):
        cell = 2.0 ** l
        # snapped origin: camera res/2 cells from the origin, on a cell boundary (closed form)
        assert params["level"][l] == (math.floor((cx - 32 * cell) / cell) * cell,
                                      math.floor((cz - 32 * cell) / cell) * cell, cell)
        assert params["offset"][l] == (0, 0)
        assert np.array_equal(u["data"], logical[l])


def _texture_after(vx, world, moves, wrap_fix):
    st = vx.ClipmapState(levels=LEVELS, res=RES, wrap_fix=wrap_fix)
    h = np.zeros((LEVELS, RES, RES), np.float32)
    clip = None
    for x, z, t in moves:
        clip, _, _ = product_step(vx, world, st, h, x, z, t)
    return clip, h


@pytest.mark.parametrize("axis", ["x", "z"])
def test_single_axis_walk_keeps_a_full_rebuild(vx, world, axis):
    # Along one axis the other axis' offset stays 0, where the reference's strip
    # placement and the sampler's storage mapping agree: incremental == rebuild.
    rng = np.random.default_rng(7 if axis == "x" else 8)
    x, z = 50.2, 75.9
    moves = [(x, z, 0.0)]
    for k in range(1, 40):
        d = float(rng.uniform(-9.0, 9.0))
        x, z = (x + d, z) if axis == "x" else (x, z + d)
        moves.append((x, z, 0.1 * k))
        clip, h = _texture_after(vx, world, moves, wrap_fix=False)
        _, want = full_logical(vx, world, x, z)
        for l in range(LEVELS):
            other = clip.offset[l][1] if axis == "x" else clip.offset[l][0]
            assert other == 0
            assert np.array_equal(vx.clip_logical(clip, h, l), want[l]), (k, l)


def test_two_axis_walk_with_wrap_fix_keeps_a_full_rebuild(vx, world):
    st = vx.ClipmapState(levels=LEVELS, res=RES, wrap_fix=True)
    h = np.zeros((LEVELS, RES, RES), np.float32)
    for k, (x, z, t) in enumerate(walk(23)):
        clip, _, _ = product_step(vx, world, st, h, x, z, t)
        _, want = full_logical(vx, world, x, z)
        for l in range(LEVELS):
            assert np.array_equal(vx.clip_logical(clip, h, l), want[l]), (k, l)


def test_reference_strip_placement_is_rotated_by_the_other_offset(vx, world):
    # +1 cell in x, then +1 cell in z: the new row (logical z = res-1) is written
    # from storage x = 0 in logical order, so the sampler, reading logical x at
    # storage (x + off_x) mod res, sees true column x + off_x there
    # (clipmap.rs:470-471 vs clipmap.wgsl:42-47).
    x0, z0 = 10.5, 20.5
    moves = [(x0, z0, 0.0), (x0 + 1.0, z0, 0.1), (x0 + 1.0, z0 + 1.0, 0.2)]
    clip, h = _texture_after(vx, world, moves, wrap_fix=False)
    _, want = full_logical(vx, world, x0 + 1.0, z0 + 1.0, levels=1)
    got = vx.clip_logical(clip, h, 0)
    offx, offz = clip.offset[0][0], clip.offset[0][1]
    assert (offx, offz) == (1, 1)
    assert np.array_equal(got[:RES - 1], want[0][:RES - 1])
    assert np.array_equal(got[RES - 1], want[0][RES - 1][(np.arange(RES) + offx) % RES])
    assert not np.array_equal(got[RES - 1], want[0][RES - 1])


def test_big_jump_rebuilds_the_level(vx, world):
    st = vx.ClipmapState(levels=LEVELS, res=RES)
    st.update(0.0, 0.0, 0.0)
    st.update(3.0, 2.0, 0.1)
    # 100 m: >= 64 cells at level 0 (1 m), 50 cells at level 1, 25 at level 2
    clip, rects = st.update(103.0, 2.0, 0.2)
    assert rects[0].as_tuple()[:7] == (0, 0, 0, RES, RES, 0, 0)
    assert tuple(clip.offset[0][:2]) == (0, 0)
    assert [r.level for r in rects[1:]] == [1, 1, 2] or [r.level for r in rects[1:]] == [1, 2]
    assert all(r.w < RES for r in rects[1:])


def test_cadence_gate_and_edge_guard(vx, world):
    st = vx.ClipmapState(levels=LEVELS, res=RES, min_update_interval_s=1.0)
    c0, r0 = st.update(0.5, 0.5, 0.0)
    assert len(r0) == LEVELS
    # a one-cell move inside the interval: no update, committed uniform unchanged
    c1, r1 = st.update(1.5, 0.5, 0.5)
    assert r1 == [] and bytes(c1) == bytes(c0)
    # after the interval the pending move lands
    c2, r2 = st.update(1.5, 0.5, 1.2)
    assert [(r.level, r.w, r.h) for r in r2] == [(0, 1, RES)]
    assert c2.level[0][0] == c0.level[0][0] + 1.0
    # the camera within 16 texels of level 0's committed edge forces an update
    # inside the interval (the camera sits at texel 32; 17 cells puts it at 49)
    c3, r3 = st.update(1.5 + 17.0, 0.5, 1.3)
    assert r3 and r3[0].level == 0 and r3[0].w == 17
    c4, r4 = st.update(1.5 + 17.0 + 1.0, 0.5, 1.4)
    assert r4 == []


def test_rejects_small_rect_buffer(vx):
    import ctypes as C
    st = vx.ClipmapState(levels=4, res=RES)
    clip = vx.Clipmap()
    rects = (vx.ClipRect * 15)()
    n = C.c_uint32()
    assert vx.LIB.vx_clipmap_update(st._h, 0.0, 0.0, 0.0, C.byref(clip), rects, 15, C.byref(n)) != 0
    assert "rectangles" in vx.LIB.vx_last_error().decode()
    # nothing changed: the next update is still the first touch
    _, r = st.update(0.0, 0.0, 0.0)
    assert len(r) == 4 and all(x.w == RES and x.h == RES for x in r)


@pytest.mark.parametrize("cell", [0.0, -1.0, float("nan"), float("inf")])
def test_build_rejects_bad_cell(vx, world, cell):
    with pytest.raises(vx.VoxhipError, match="bad argument"):
        vx.clipmap_build(world, 0.0, 0.0, levels=2, res=RES, base_cell_m=cell)
