"""GPU: the incremental clipmap with device-generated heights.

vx_write_clipmap_rects computes the texels of the planner's rectangles with a
HIP kernel straight into the resident clip array.  Checked bit for bit against
the host fill (vx_clipmap_fill_rects + vx_write_clipmap) over camera walks with
strips, wraps and full rebuilds, in both strip placements; then the full
pipeline along the seeded flight with the clipmap updated every frame against
the oracle fed the same texture and uniform.
"""
from __future__ import annotations

import json

import numpy as np
import pytest

from conftest import host_world
from parity import COLOR_ABS_TOL, DEPTH_REL_TOL, make_context, summary

pytestmark = pytest.mark.gpu

LEVELS, RES = 5, 256


def _moves(sc):
    """The flight's camera xz with time, then strides and two teleports."""
    out = [(float(p.pos[0]), float(p.pos[2]), t) for p, t in sc.poses]
    x, z, t = out[-1]
    rng = np.random.default_rng(17)
    for k in range(30):
        if k in (10, 22):
            x += 700.0 if k == 10 else -2100.0
            z -= 350.0
        else:
            x += float(rng.uniform(-40, 40))
            z += float(rng.uniform(-40, 40))
        out.append((x, z, t + 0.1 * (k + 1)))
    return out


@pytest.mark.parametrize("wrap_fix", [False, True])
def test_device_fill_matches_host_fill(gpu, wrap_fix):
    sc, hw = host_world("c1")
    ctx_dev = gpu.Context(node_capacity=16, chunk_capacity=1, grid_capacity=1, clip_levels=LEVELS, clip_res=RES)
    ctx_host = gpu.Context(node_capacity=16, chunk_capacity=1, grid_capacity=1, clip_levels=LEVELS, clip_res=RES)
    st = gpu.ClipmapState(levels=LEVELS, res=RES, wrap_fix=wrap_fix)
    mirror = np.zeros((LEVELS, RES, RES), np.float32)
    n_rects = n_full = 0
    for k, (x, z, t) in enumerate(_moves(sc)):
        clip, rects = st.update(x, z, t)
        data = gpu.clipmap_fill_rects(sc.params, rects)
        gpu.apply_clip_rects(mirror, rects, data)
        ctx_dev.write_clipmap_rects(clip, sc.params, rects)
        ctx_host.write_clipmap_patches(clip, rects, data)
        n_rects += len(rects)
        n_full += sum(r.w == RES and r.h == RES for r in rects)
        if rects or k % 16 == 0:
            got = ctx_dev.read_clipmap()
            assert np.array_equal(got.view(np.uint32), mirror.view(np.uint32)), (k, [r.as_tuple() for r in rects])
            assert np.array_equal(ctx_host.read_clipmap().view(np.uint32), mirror.view(np.uint32)), k
    print(f"[clipmap] wrap_fix={wrap_fix}: {n_rects} rectangles, {n_full} full levels", flush=True)
    # first touch: 5 levels; +700 m: levels 0-1 (>= 256 cells); -2100 m: levels 0-3
    assert n_full == LEVELS + 2 + 4
    ctx_dev.close()
    ctx_host.close()


def test_write_clipmap_rects_rejects_bad_rectangles(gpu):
    sc, _ = host_world("c1")
    ctx = gpu.Context(node_capacity=16, chunk_capacity=1, grid_capacity=1, clip_levels=2, clip_res=64)
    st = gpu.ClipmapState(levels=2, res=64)
    clip, rects = st.update(5.0, 5.0, 0.0)
    bad = gpu.ClipRect.from_buffer_copy(rects[0])
    bad.x = 1  # 1 + 64 > 64
    with pytest.raises(gpu.VoxhipError, match="out of range"):
        ctx.write_clipmap_rects(clip, sc.params, [bad])
    bad = gpu.ClipRect.from_buffer_copy(rects[0])
    bad.level = 2
    with pytest.raises(gpu.VoxhipError, match="out of range"):
        ctx.write_clipmap_rects(clip, sc.params, [bad])
    with pytest.raises(gpu.VoxhipError, match="rectangles"):
        ctx.write_clipmap_rects(clip, sc.params, rects * 40)
    # host patches whose x + w wraps in 32 bits are rejected before any copy
    bad = gpu.ClipRect.from_buffer_copy(rects[0])
    bad.x, bad.w, bad.h = 0xFFFFFFF0, 32, 1
    with pytest.raises(gpu.VoxhipError, match="out of range"):
        ctx.write_clipmap_patches(clip, [bad], [np.zeros(32, np.float32)])
    # the context is still usable
    ctx.write_clipmap_rects(clip, sc.params, rects)
    h = ctx.read_clipmap()
    assert np.array_equal(h[0], gpu.clipmap_fill_rects(sc.params, rects)[0])
    ctx.close()


def test_flight_with_incremental_clipmap_matches_oracle(gpu):
    """The bench world along the seeded flight with the clipmap updated every
    frame (the reference's frame order: clipmap update, uploads, passes), the
    reference's strip placement, full pipeline at 480x270 for frames 60..62,
    against the oracle fed the same texture and uniform."""
    from oracle import oracle as O

    sc, hw = host_world("c1")
    res = (480, 270)
    ctx = make_context(gpu, hw)
    st = gpu.ClipmapState(levels=LEVELS, res=RES)
    mirror = np.zeros((LEVELS, RES, RES), np.float32)
    view = gpu.View(ctx, res, res)
    fr = O.Frame(*res)
    prev = None
    moved = 0
    for f in range(63):
        pose, t_s = sc.poses[f]
        clip, rects = st.update(float(pose.pos[0]), float(pose.pos[2]), t_s)
        gpu.apply_clip_rects(mirror, rects, gpu.clipmap_fill_rects(sc.params, rects))
        ctx.write_clipmap_rects(clip, sc.params, rects)
        moved += f > 0 and bool(rects)
        if f < 60:
            continue
        cam = gpu.camera_build(pose, res, res, f, time_s=t_s, wind_strength=2.0, grid_origin=sc.grid_origin,
                               grid_dims=sc.grid_dims, chunk_count=hw.chunk_count, prev_view_proj=prev, jitter=True)
        prev = np.array(cam.view_proj, np.float32)
        view.write_camera(cam)
        view.render(gpu.PASS_FRAME | gpu.PASS_WRITE_HITS)
        w = O.World.from_host(hw, clip, mirror.copy())
        O.frame(w, cam, fr, gpu.HIT_DTYPE)
        s = summary(view.read(gpu.OUT_HITS), fr.hits, view.read(gpu.OUT_DEPTH), fr.depth, view.read(gpu.OUT_COLOR),
                    fr.color)
        s["output_max_abs"] = float(np.max(np.abs(view.read(gpu.OUT_OUTPUT) - fr.output)))
        s["kinds"] = np.bincount(fr.hits["kind"].ravel(), minlength=3).tolist()
        s["offsets"] = [list(clip.offset[l][:2]) for l in range(LEVELS)]
        print(f"[parity] c1 incremental clipmap frame {f}: {json.dumps(s)}", flush=True)
        assert s["identity_mismatches"] == 0, s
        assert s["depth_max_rel"] <= DEPTH_REL_TOL, s
        assert s["color_max_abs"] <= COLOR_ABS_TOL, s
        assert s["output_max_abs"] <= COLOR_ABS_TOL, s
        assert s["kinds"][2] > 0  # heightfield pixels read the updated texture
    assert moved > 0
    view.close()
    ctx.close()


def test_wrap_fix_texture_renders_as_a_full_build(gpu):
    """With VX_CLIP_WRAP_FIX the incrementally updated texture holds a full
    rebuild's heights under different offsets: the rendered frame is bit-identical
    to one rendered from vx_clipmap_build at the same camera."""
    from voxhip import scenes

    sc, hw = host_world("c1")
    res = (320, 180)
    ctx_inc = make_context(gpu, hw)
    ctx_full = make_context(gpu, hw)
    st = gpu.ClipmapState(levels=LEVELS, res=RES, wrap_fix=True)
    frame = 90
    for f in range(frame + 1):
        pose, t_s = sc.poses[f]
        clip, rects = st.update(float(pose.pos[0]), float(pose.pos[2]), t_s)
        ctx_inc.write_clipmap_rects(clip, sc.params, rects)
    full_clip, full_h = scenes.scene_clipmap(sc, frame)
    ctx_full.write_clipmap(full_clip, full_h)
    assert any(clip.offset[l][0] != full_clip.offset[l][0] for l in range(LEVELS))
    outs = []
    for ctx in (ctx_inc, ctx_full):
        view = gpu.View(ctx, res, res)
        cam = scenes.frame_camera(sc, hw, frame, render=res)
        view.write_camera(cam)
        view.render(gpu.PASS_FRAME | gpu.PASS_WRITE_HITS)
        outs.append([view.read(k) for k in (gpu.OUT_HITS, gpu.OUT_DEPTH, gpu.OUT_COLOR, gpu.OUT_OUTPUT)])
        view.close()
    for a, b in zip(*outs):
        assert np.array_equal(np.ascontiguousarray(a).view(np.uint8), np.ascontiguousarray(b).view(np.uint8))
    ctx_inc.close()
    ctx_full.close()
