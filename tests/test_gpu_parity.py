"""GPU parity: the HIP passes against the oracle on the same seeded inputs.

hit identity bit-exact, depth within 1e-5 relative, colour
within 1e-3 absolute (parity.py).  All runs go through the C ABI.
"""
from __future__ import annotations

import json
import math

import numpy as np
import pytest

from conftest import host_world
from parity import COLOR_ABS_TOL, DEPTH_REL_TOL, make_context, summary

pytestmark = pytest.mark.gpu


def _oracle():
    from oracle import oracle as O

    return O


def _report(name, s):
    print(f"[parity] {name}: {json.dumps(s)}", flush=True)


def _assert_parity(s):
    assert s["identity_mismatches"] == 0, s
    assert s["depth_max_rel"] <= DEPTH_REL_TOL, s
    assert s["color_max_abs"] <= COLOR_ABS_TOL, s


@pytest.mark.parametrize("wind", [0.0, 2.0])
def test_primary_c0_first_frame(gpu, wind):
    """Config 0: plain primary pass, 640x360, fresh history, fixed seeded pose."""
    O = _oracle()
    from voxhip import scenes

    sc, hw = host_world("c0")
    clip, heights = scenes.scene_clipmap(sc)
    cam = scenes.frame_camera(sc, hw, 0, wind_strength=wind)
    ctx = make_context(gpu, hw)
    ctx.write_clipmap(clip, heights)
    view = gpu.View(ctx, sc.present, sc.render)
    view.write_camera(cam)
    view.render(gpu.PASS_PRIMARY | gpu.PASS_WRITE_HITS | gpu.PASS_END_FRAME)
    hits = view.read(gpu.OUT_HITS)
    depth = view.read(gpu.OUT_DEPTH)
    color = view.read(gpu.OUT_COLOR)
    local = view.read(gpu.OUT_LOCAL)
    phist = view.read(gpu.OUT_PRIMARY_HIST)
    shadow = view.read(gpu.OUT_SHADOW_HIST)

    w = O.World.from_host(hw, clip, heights)
    fr = O.Frame(*sc.render)
    O.primary(w, cam, fr, gpu.HIT_DTYPE)
    s = summary(hits, fr.hits, depth, fr.depth, color, fr.color)
    s["local_max_abs"] = float(np.max(np.abs(local - fr.local)))
    s["hist_bit_exact"] = bool(np.array_equal(phist.view(np.uint32), fr.primary_hist.view(np.uint32)))
    s["shadow_max_abs"] = float(np.max(np.abs(shadow - fr.shadow)))
    s["kinds"] = np.bincount(fr.hits["kind"].ravel(), minlength=3).tolist()
    _report(f"c0 primary wind={wind}", s)
    _assert_parity(s)
    assert s["local_max_abs"] <= COLOR_ABS_TOL
    assert s["shadow_max_abs"] <= COLOR_ABS_TOL
    # Every ray class appears, so the comparison covers voxel, heightfield and sky paths.
    assert min(s["kinds"]) > 0


@pytest.mark.parametrize("yaw,pitch", [(0.0, 0.0), (math.pi / 2, -math.pi / 6), (math.pi / 4, -math.pi / 4)])
def test_primary_camera_on_chunk_corner(gpu, yaw, pitch):
    """The camera exactly on a chunk corner (x and z on chunk faces, y on a
    voxel face), looking along an axis, an axis-aligned 30-degree dive and a
    diagonal: every ray starts on the ties of the inward-biased chunk coordinate
    (common.wgsl:543-554) and the probe rays' chunk DDA; 256x144 primary pass
    against the oracle."""
    O = _oracle()
    from voxhip import scenes

    sc, hw = host_world("c0")
    pose0, t0 = sc.poses[0]
    f = np.float32
    cx, cz = (round(float(pose0.pos[i]) / 6.4) for i in (0, 2))
    pos = (float(f(cx) * f(6.4)), float(f(round(float(pose0.pos[1]) / 0.1)) * f(0.1)), float(f(cz) * f(6.4)))
    res = (256, 144)
    cam = gpu.camera_build(gpu.make_pose(pos, yaw, pitch), res, res, 0, time_s=t0, grid_origin=sc.grid_origin,
                           grid_dims=sc.grid_dims, chunk_count=hw.chunk_count)
    clip, heights = scenes.scene_clipmap(sc)
    ctx = make_context(gpu, hw)
    ctx.write_clipmap(clip, heights)
    view = gpu.View(ctx, res, res)
    view.write_camera(cam)
    view.render(gpu.PASS_PRIMARY | gpu.PASS_WRITE_HITS | gpu.PASS_END_FRAME)
    w = O.World.from_host(hw, clip, heights)
    fr = O.Frame(*res)
    O.primary(w, cam, fr, gpu.HIT_DTYPE)
    s = summary(view.read(gpu.OUT_HITS), fr.hits, view.read(gpu.OUT_DEPTH), fr.depth, view.read(gpu.OUT_COLOR),
                fr.color)
    s["hist_bit_exact"] = bool(np.array_equal(view.read(gpu.OUT_PRIMARY_HIST).view(np.uint32),
                                              fr.primary_hist.view(np.uint32)))
    s["kinds"] = np.bincount(fr.hits["kind"].ravel(), minlength=3).tolist()
    _report(f"chunk-corner camera {pos} yaw={yaw:.3f} pitch={pitch:.3f}", s)
    _assert_parity(s)
    assert s["hist_bit_exact"]
    assert s["kinds"][1] > 0  # voxel hits
    view.close()
    ctx.close()


def test_light_gather_after_reupload(gpu):
    """The light gather skips its march where no chunk within reach holds a
    LIGHT leaf; the per-chunk flags it reads are derived on the device after
    each upload.  Render the unlit c0 world, then upload the same world with
    LIGHT voxels placed in view into the same context and render again: both
    frames must match the oracle, and the second must receive local light."""
    O = _oracle()
    from voxhip import scenes

    sc, hw = host_world("c0")
    _, hw_lit = host_world("c0_lit")
    clip, heights = scenes.scene_clipmap(sc)
    ctx = gpu.Context(node_capacity=max(len(hw.nodes), len(hw_lit.nodes)), chunk_capacity=hw.chunk_count,
                      grid_capacity=len(hw.grid), clip_levels=5, clip_res=256)
    ctx.write_clipmap(clip, heights)
    view = gpu.View(ctx, sc.present, sc.render)
    for name, world in (("unlit", hw), ("lit", hw_lit)):
        ctx.upload_world(world)
        cam = scenes.frame_camera(sc, world, 0)
        view.reset()
        view.write_camera(cam)
        view.render(gpu.PASS_PRIMARY | gpu.PASS_WRITE_HITS | gpu.PASS_END_FRAME)
        w = O.World.from_host(world, clip, heights)
        fr = O.Frame(*sc.render)
        O.primary(w, cam, fr, gpu.HIT_DTYPE)
        s = summary(view.read(gpu.OUT_HITS), fr.hits, view.read(gpu.OUT_DEPTH), fr.depth, view.read(gpu.OUT_COLOR), fr.color)
        local = view.read(gpu.OUT_LOCAL)
        s["local_max_abs"] = float(np.max(np.abs(local - fr.local)))
        s["lit_pixels"] = int(np.sum(np.any(fr.local[..., :3] > 0.0, axis=-1)))
        _report(f"c0 light gather {name}", s)
        _assert_parity(s)
        assert s["local_max_abs"] <= COLOR_ABS_TOL
        assert (s["lit_pixels"] > 0) == (name == "lit")


@pytest.mark.parametrize("fog", [0.002, 0.0])
def test_full_pipeline_multiframe(gpu, fog):
    """Four full frames (primary, local TAA, godray, composite, composite TAA)
    with a moving camera, so history reprojection, anchors and the 1/8 shadow
    refresh feed back frame to frame on both sides."""
    O = _oracle()
    from voxhip import scenes

    sc, hw = host_world("c0")
    clip, heights = scenes.scene_clipmap(sc)
    res = (316, 180)  # not multiples of 8: the last tile row and column are partial
    ctx = make_context(gpu, hw)
    ctx.write_clipmap(clip, heights)
    view = gpu.View(ctx, res, res)
    w = O.World.from_host(hw, clip, heights)
    fr = O.Frame(*res)
    pose0, _ = sc.poses[0]
    prev = None
    for f in range(4):
        pose = gpu.make_pose((pose0.pos[0] + 0.05 * f, pose0.pos[1], pose0.pos[2] - 0.03 * f), pose0.yaw + 0.01 * f,
                             pose0.pitch, 70.0)
        cam = gpu.camera_build(pose, res, res, f, time_s=f / 60.0, wind_strength=2.0, fog_density=fog,
                               grid_origin=sc.grid_origin, grid_dims=sc.grid_dims, chunk_count=hw.chunk_count,
                               prev_view_proj=prev, jitter=True)
        prev = np.array(cam.view_proj, np.float32)
        view.write_camera(cam)
        view.render(gpu.PASS_FRAME | gpu.PASS_WRITE_HITS)
        O.frame(w, cam, fr, gpu.HIT_DTYPE)
        s = summary(view.read(gpu.OUT_HITS), fr.hits, view.read(gpu.OUT_DEPTH), fr.depth, view.read(gpu.OUT_COLOR),
                    fr.color)
        out = view.read(gpu.OUT_OUTPUT)
        s["output_max_abs"] = float(np.max(np.abs(out - fr.output)))
        s["godray_max_abs"] = float(np.max(np.abs(view.read(gpu.OUT_GODRAY) - fr.godray)))
        s["local_hist_max_abs"] = float(np.max(np.abs(view.read(gpu.OUT_LOCAL_HIST) - fr.local_hist)))
        _report(f"pipeline frame {f} fog={fog}", s)
        _assert_parity(s)
        if fog == 0.0:  # no fog: the godray march is off and its output is exactly zero
            assert not view.read(gpu.OUT_GODRAY)[..., :3].any()
        assert s["output_max_abs"] <= COLOR_ABS_TOL
        assert s["godray_max_abs"] <= COLOR_ABS_TOL
        assert s["local_hist_max_abs"] <= COLOR_ABS_TOL


@pytest.mark.parametrize("fix", [False, True])
def test_godray_macro_binding_modes(gpu, fix):
    """Godray with the reference's rope-buffer macro binding (default) and the fixed binding."""
    O = _oracle()
    from voxhip import scenes

    sc, hw = host_world("c0")
    clip, heights = scenes.scene_clipmap(sc)
    res = (256, 144)
    ctx = make_context(gpu, hw, godray_macro_fix=fix)
    ctx.write_clipmap(clip, heights)
    view = gpu.View(ctx, res, res)
    w = O.World.from_host(hw, clip, heights)
    fr = O.Frame(*res)
    pose, _ = sc.poses[0]
    cam = gpu.camera_build(pose, res, res, 5, wind_strength=0.0, grid_origin=sc.grid_origin, grid_dims=sc.grid_dims,
                           chunk_count=hw.chunk_count)
    view.write_camera(cam)
    view.render(gpu.PASS_PRIMARY | gpu.PASS_GODRAY | gpu.PASS_END_FRAME)
    O.primary(w, cam, fr)
    O.godray(w, cam, fr, macro_compat=not fix)
    err = float(np.max(np.abs(view.read(gpu.OUT_GODRAY) - fr.godray)))
    print(f"[parity] godray fix={fix}: max_abs={err}", flush=True)
    assert err <= COLOR_ABS_TOL


def test_flight_c1_world(gpu):
    """The bench world (16x4x16 chunks, 5-level clipmap) along the seeded
    camera flight, full pipeline, three frames at 640x360: the same world,
    grass density and camera path as the 4K benchmark, at a size the oracle
    renders in seconds."""
    O = _oracle()
    from voxhip import scenes

    sc, hw = host_world("c1")
    clip, heights = scenes.scene_clipmap(sc, 0)
    res = (640, 360)
    ctx = make_context(gpu, hw)
    ctx.write_clipmap(clip, heights)
    view = gpu.View(ctx, res, res)
    w = O.World.from_host(hw, clip, heights)
    fr = O.Frame(*res)
    prev = None
    for f in range(3):
        frame = 30 + f
        pose, t_s = sc.poses[frame]
        cam = gpu.camera_build(pose, res, res, frame, time_s=t_s, wind_strength=2.0, grid_origin=sc.grid_origin,
                               grid_dims=sc.grid_dims, chunk_count=hw.chunk_count, prev_view_proj=prev, jitter=True)
        prev = np.array(cam.view_proj, np.float32)
        view.write_camera(cam)
        view.render(gpu.PASS_FRAME | gpu.PASS_WRITE_HITS)
        O.frame(w, cam, fr, gpu.HIT_DTYPE)
        s = summary(view.read(gpu.OUT_HITS), fr.hits, view.read(gpu.OUT_DEPTH), fr.depth, view.read(gpu.OUT_COLOR),
                    fr.color)
        s["output_max_abs"] = float(np.max(np.abs(view.read(gpu.OUT_OUTPUT) - fr.output)))
        s["kinds"] = np.bincount(fr.hits["kind"].ravel(), minlength=3).tolist()
        s["grass_hits"] = int(np.sum((fr.hits["kind"] == 1) & (fr.hits["mat"] == gpu.MAT_GRASS)))
        _report(f"c1 flight frame {frame}", s)
        _assert_parity(s)
        assert s["output_max_abs"] <= COLOR_ABS_TOL


def test_counted_loads_match_oracle(gpu):
    """The counting form of the primary and godray passes (the roofline's
    algorithmic bytes) against the oracle's own load counts for the same frame:
    every counter equal; the per-ray byte totals sum to the launch's bytes."""
    O = _oracle()
    from voxhip import scenes

    sc, hw = host_world("c0")
    clip, heights = scenes.scene_clipmap(sc)
    cam = scenes.frame_camera(sc, hw, 0, wind_strength=2.0)
    ctx = make_context(gpu, hw)
    ctx.write_clipmap(clip, heights)
    view = gpu.View(ctx, sc.present, sc.render)
    with pytest.raises(gpu.VoxhipError, match="counting"):
        view.read(gpu.OUT_RAY_BYTES)
    view.write_camera(cam)
    view.counters(reset=True)
    view.render(gpu.PASS_PRIMARY | gpu.PASS_COUNT)
    c_prim = view.counters(reset=True)
    ray_bytes = view.read(gpu.OUT_RAY_BYTES).astype(np.float64)
    view.render(gpu.PASS_GODRAY | gpu.PASS_COUNT | gpu.PASS_END_FRAME)
    c_god = view.counters(reset=True)

    w = O.World.from_host(hw, clip, heights)
    fr = O.Frame(*sc.render)
    oc = O.Counts()
    O.primary(w, cam, fr, gpu.HIT_DTYPE, counts=oc)
    og = O.Counts()
    O.godray(w, cam, fr, counts=og)
    _report("counted loads", {"device_primary": c_prim, "oracle_primary": oc.as_dict(), "device_godray": c_god,
                              "oracle_godray": og.as_dict()})
    for n in gpu.COUNT_NAMES:
        assert c_prim[n] == oc.as_dict()[n], n
        assert c_god[n] == og.as_dict()[n], n
    # 640x360 is a whole number of 8x8 tiles: the tile-gather shares add up exactly
    assert ray_bytes.sum() == gpu.algorithmic_bytes(c_prim)
    assert ray_bytes.min() > 0


def test_profile_counters_match_oracle(gpu):
    """The reference's five profile counters (ray_main.wgsl:32-74), raised while
    the camera's profile_flags & 1: through vx_render and through a batch, equal
    to the oracle's over three C1 flight frames at 640x360 (so windowed and
    full traces, grass refinements, traced and reused shadows all occur); a
    view in the same batch with the flag off stays at zero; reset clears them."""
    O = _oracle()
    from voxhip import scenes

    sc, hw = host_world("c1")
    clip, heights = scenes.scene_clipmap(sc, 0)
    res = (640, 360)
    ctx = make_context(gpu, hw)
    ctx.write_clipmap(clip, heights)
    single = gpu.View(ctx, res, res)
    in_batch, quiet = gpu.View(ctx, res, res), gpu.View(ctx, (320, 180), (320, 180))
    batch = gpu.Batch(ctx, [in_batch, quiet])
    w = O.World.from_host(hw, clip, heights)
    fr = O.Frame(*res)
    prev = None
    for f in range(3):
        frame = 60 + f
        cam = scenes.frame_camera(sc, hw, frame, prev_view_proj=prev, jitter=True, render=res)
        prev = np.array(cam.view_proj, np.float32)
        cam.profile_flags = 1
        oc = O.Counts()
        O.primary(w, cam, fr, counts=oc)   # the oracle's pipeline, with the primary pass counted
        O.local_taa(fr)
        O.godray(w, cam, fr)
        O.composite(cam, fr)
        O.composite_taa(cam, fr)
        ref = oc.profile()
        single.write_camera(cam)
        single.render(gpu.PASS_FRAME)
        in_batch.write_camera(cam)
        quiet.write_camera(scenes.frame_camera(sc, hw, frame, render=(320, 180)))   # profile_flags 0
        batch.render(gpu.PASS_FRAME)
        got_single, got_batch, got_quiet = single.profile(), in_batch.profile(), quiet.profile()
        print(f"[parity] profile counters frame {frame}: oracle {ref}, vx_render {got_single}, batch {got_batch}",
              flush=True)
        assert got_single == ref and got_batch == ref, (ref, got_single, got_batch)
        assert all(v == 0 for v in got_quiet.values()), got_quiet
        assert all(v > 0 for v in ref.values()), ref
        assert np.array_equal(single.read(gpu.OUT_OUTPUT).view(np.uint32), fr.output.view(np.uint32))
    assert all(v == 0 for v in single.profile().values())   # the reads above reset them
    single.render(gpu.PASS_PRIMARY)   # the same camera again, no reset in between: the counters add up
    single.render(gpu.PASS_PRIMARY)
    twice = single.profile(reset=False)
    assert twice == single.profile() and all(v > 0 for v in twice.values())
    batch.close()
