"""The batching, ray-query and serving workloads at their benchmark sizes on
the GPU, each tied to the oracle, plus two parity cases the smaller tests do
not reach:

  * the 24-candidate truncation of the tile gather (a low camera looking
    diagonally over the 16x4x16 world: tiles whose probe rays cross more than
    24 non-empty chunks);
  * render scale 0.5 (present 640x360, render 320x180), the present<->render
    mapping of composite and composite TAA;
  * 64 views of mixed resolution (320x180 .. 3840x2160, seed 7) rendered as
    one batch equal, output for output, the same views rendered one per
    launch, and a 320x180 view of the batch equals the oracle;
  * 256 ray batches (log-uniform 1e3..1e6 rays, seed 11) over the 32x4x32
    world as one fused launch equal one launch per batch, and a seeded 1 %
    sample plus the boundary rays equal the oracle;
  * Poisson serving (400 req/s) with a concurrent edit stream (1000 single-
    voxel edits/s) on the 32x4x32 world: every served ray batch equals the
    oracle on the world as it stood (the edits are confined to chunks the ray
    batches cannot reach), and after the edits are applied a served frame and
    ray batch equal the oracle on a world rebuilt with every edit."""
from __future__ import annotations

import json
import math
import threading
import time

import numpy as np
import pytest

import worlds
from conftest import host_world
from parity import COLOR_ABS_TOL, DEPTH_REL_TOL, make_context, summary

pytestmark = pytest.mark.gpu


def _oracle():
    from oracle import oracle as O

    return O


def _assert_parity(name, s):
    print(f"[configs] {name}: {json.dumps(s)}", flush=True)
    assert s["identity_mismatches"] == 0, s
    assert s["depth_max_rel"] <= DEPTH_REL_TOL, s
    assert s["color_max_abs"] <= COLOR_ABS_TOL, s


def _scene_cam(gpu, grid_origin, grid_dims, chunk_count):
    return gpu.camera_build(gpu.make_pose((1.0, 1.0, 1.0), 0.0, 0.0), (64, 64), (64, 64), 0, grid_origin=grid_origin,
                            grid_dims=grid_dims, chunk_count=chunk_count)


def _same_rows(a, b):
    """Rows (hit records) whose bytes differ."""
    n = len(a)
    return np.nonzero((a.view(np.uint8).reshape(n, -1) != b.view(np.uint8).reshape(n, -1)).any(axis=1))[0]


# ------------------------------------------------------------ truncation
def test_tile_truncation_reached_and_exact(gpu):
    """H1: only the first 24 of a tile's sorted candidate chunks are traced.
    A camera 3 m above the ground looking diagonally across the world at the
    horizon gives tiles with 25-40 raw candidates; two full frames (the second
    with history) match the oracle bit for bit."""
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
    for f in range(2):
        x = 10.0 + 0.08 * f
        pose = gpu.make_pose((x, scenes.surface_m(sc.params, 10.0, 10.0) + 3.0, 10.0 + 0.08 * f), math.pi / 4,
                             math.radians(-2.0), 70.0)
        cam = gpu.camera_build(pose, res, res, f, time_s=f / 60.0, wind_strength=2.0, grid_origin=sc.grid_origin,
                               grid_dims=sc.grid_dims, chunk_count=hw.chunk_count, prev_view_proj=prev, jitter=True)
        prev = np.array(cam.view_proj, np.float32)
        raw = O.tile_raw_counts(w, cam, *res)
        assert (raw > 24).sum() >= 50, int((raw > 24).sum())   # the truncation applies to these tiles
        view.write_camera(cam)
        view.render(gpu.PASS_FRAME | gpu.PASS_WRITE_HITS)
        O.frame(w, cam, fr, gpu.HIT_DTYPE)
        s = summary(view.read(gpu.OUT_HITS), fr.hits, view.read(gpu.OUT_DEPTH), fr.depth, view.read(gpu.OUT_COLOR),
                    fr.color)
        s["tiles_over_24"] = int((raw > 24).sum())
        s["max_raw_candidates"] = int(raw.max())
        s["output_max_abs"] = float(np.max(np.abs(view.read(gpu.OUT_OUTPUT) - fr.output)))
        _assert_parity(f"truncation frame {f}", s)
        assert s["output_max_abs"] <= COLOR_ABS_TOL


# ------------------------------------------------------------ render scale
def test_render_scale_half_pipeline(gpu):
    """The reference app renders at half scale (app/config.rs:16): present
    640x360 from a 320x180 render, four frames of the C1 flight, full
    pipeline.  Composite maps present pixels to render pixels and composite
    TAA reprojects in present space (common.wgsl:664-690, taa.wgsl:83-110)."""
    O = _oracle()
    from voxhip import scenes

    sc, hw = host_world("c1")
    clip, heights = scenes.scene_clipmap(sc, 0)
    present, render = (640, 360), (320, 180)
    ctx = make_context(gpu, hw)
    ctx.write_clipmap(clip, heights)
    view = gpu.View(ctx, present, render)
    w = O.World.from_host(hw, clip, heights)
    fr = O.Frame(*render, *present)
    prev = None
    for f in range(4):
        frame = 40 + f
        cam = scenes.frame_camera(sc, hw, frame, prev_view_proj=prev, jitter=True, render=render, present=present)
        prev = np.array(cam.view_proj, np.float32)
        view.write_camera(cam)
        view.render(gpu.PASS_FRAME | gpu.PASS_WRITE_HITS)
        O.frame(w, cam, fr, gpu.HIT_DTYPE)
        s = summary(view.read(gpu.OUT_HITS), fr.hits, view.read(gpu.OUT_DEPTH), fr.depth, view.read(gpu.OUT_COLOR),
                    fr.color)
        pre, out = view.read(gpu.OUT_PRE_TAA), view.read(gpu.OUT_OUTPUT)
        assert pre.shape == (360, 640, 4) and out.shape == (360, 640, 4)
        s["pre_taa_max_abs"] = float(np.max(np.abs(pre - fr.pre_taa)))
        s["output_max_abs"] = float(np.max(np.abs(out - fr.output)))
        s["output_bit_exact"] = bool(np.array_equal(out.view(np.uint32), fr.output.view(np.uint32)))
        _assert_parity(f"render scale 0.5 frame {frame}", s)
        assert s["pre_taa_max_abs"] <= COLOR_ABS_TOL and s["output_max_abs"] <= COLOR_ABS_TOL


# ------------------------------------------------------------ config 2
def _digest(a: np.ndarray) -> str:
    import xxhash

    return xxhash.xxh3_128_hexdigest(memoryview(np.ascontiguousarray(a)).cast("B"))


def test_c2_64_views_batched_equals_plain_and_oracle(gpu):
    """64 views, resolutions uniform over 320x180..3840x2160 (seed 7: 148-180 M
    pixels per frame), two frames each with history: the fused batch (the
    autotuner's configuration, graph replay) equals one launch sequence per
    view in the final output and the primary history (hit distance and anchor
    identity); the first 320x180 view equals the oracle."""
    O = _oracle()
    from voxhip import scenes

    sc, hw = host_world("c1")
    clip, heights = scenes.scene_clipmap(sc, 0)
    ctx = make_context(gpu, hw)
    ctx.write_clipmap(clip, heights)
    res, frames = scenes.batch_views(64, len(sc.poses))
    assert any(r == (3840, 2160) for r in res) and any(r == (1920, 1080) for r in res)
    steps = 2
    cams = []
    prev = [None] * len(res)
    for step in range(steps):
        cs = []
        for i, r in enumerate(res):
            c = scenes.frame_camera(sc, hw, int(frames[i]) + step, prev_view_proj=prev[i], jitter=True, render=r)
            prev[i] = np.array(c.view_proj, np.float32)
            cs.append(c)
        cams.append(cs)

    def digests(views):
        return [(_digest(v.read(gpu.OUT_OUTPUT)), _digest(v.read(gpu.OUT_PRIMARY_HIST))) for v in views]

    t0 = time.perf_counter()
    views = [gpu.View(ctx, r, r) for r in res]
    plain = []
    for step in range(steps):
        for v, c in zip(views, cams[step]):
            v.write_camera(c)
            v.render(gpu.PASS_FRAME)
        plain.append(digests(views))
    for v in views:
        v.close()
    t_plain = time.perf_counter() - t0

    views = [gpu.View(ctx, r, r) for r in res]
    batch = gpu.Batch(ctx, views)
    for v, c in zip(views, cams[0]):
        v.write_camera(c)
    tuned = batch.tune(gpu.PASS_FRAME, iters=1).as_dict()
    small = res.index((320, 180))
    O_w = O.World.from_host(hw, clip, heights)
    fr = O.Frame(320, 180)
    for step in range(steps):
        for v, c in zip(views, cams[step]):
            v.write_camera(c)
        batch.render(gpu.PASS_FRAME)
        got = digests(views)
        bad = [i for i in range(len(res)) if got[i] != plain[step][i]]
        assert not bad, (step, [(i, res[i]) for i in bad[:5]])
        O.frame(O_w, cams[step][small], fr)
        out = views[small].read(gpu.OUT_OUTPUT)
        err = float(np.max(np.abs(out - fr.output)))
        derr = float(np.max(np.abs(views[small].read(gpu.OUT_DEPTH) - fr.depth) / np.maximum(fr.depth, 1e-30)))
        print(f"[configs] c2 step {step}: {len(res)} views, {sum(w * h for w, h in res)} px, batched == plain; "
              f"view {small} vs oracle: output max abs {err:.3g}, depth max rel {derr:.3g}; tuned {tuned}; "
              f"plain leg {t_plain:.1f} s", flush=True)
        assert err <= COLOR_ABS_TOL and derr <= DEPTH_REL_TOL
    batch.close()
    for v in views:
        v.close()


# ------------------------------------------------------------ config 3
@pytest.fixture(scope="module")
def world_32x4x32(gpu):
    from voxhip import scenes

    dims = (32, 4, 32)
    hw = gpu.HostWorld(gpu.world_params(42), (0, 0, 0), dims)
    return dims, hw, scenes


def test_c3_fused_queries_32x4x32(gpu, world_32x4x32):
    O = _oracle()
    dims, hw, scenes = world_32x4x32
    origin = (0, 0, 0)
    ctx = make_context(gpu, hw, clip_levels=1)
    cam = _scene_cam(gpu, origin, dims, hw.chunk_count)
    sizes, rays = scenes.query_batches(256, origin, dims)
    n = len(rays)
    assert n > 20_000_000 and hw.chunk_count == 4096
    d_rays = gpu.DeviceArray.from_numpy(rays)
    d_fused = gpu.DeviceArray(n * gpu.HIT_DTYPE.itemsize)
    d_plain = gpu.DeviceArray(n * gpu.HIT_DTYPE.itemsize)
    gpu.query_rays_device(ctx, cam, d_rays, d_fused, n)
    ctx.sync()
    fused_ms = gpu.query_timing(ctx)
    offs = np.concatenate([[0], np.cumsum(sizes)])
    stride = gpu.HIT_DTYPE.itemsize
    for i in range(len(sizes)):
        k = int(sizes[i])
        gpu.query_rays_device(ctx, cam, d_rays.view(int(offs[i]) * 32, k * 32),
                              d_plain.view(int(offs[i]) * stride, k * stride), k)
    ctx.sync()
    fused = d_fused.download(np.empty(n, gpu.HIT_DTYPE))
    plain = d_plain.download(np.empty(n, gpu.HIT_DTYPE))
    for d in (d_rays, d_fused, d_plain):
        d.close()
    assert np.array_equal(fused.view(np.uint8), plain.view(np.uint8))
    # the oracle on a seeded 1 % sample of the rays (every batch contributes) and on the boundary rays
    sample = np.sort(np.random.default_rng(3).choice(n, n // 100, replace=False))
    clip, heights = worlds.empty_clipmap(gpu)
    w = O.World.from_host(hw, clip, heights)
    ref = O.query_rays(w, cam, rays[sample], gpu.HIT_DTYPE)
    rows = _same_rows(fused[sample], ref)
    brays = worlds.boundary_rays()
    bref = O.query_rays(w, cam, brays, gpu.HIT_DTYPE)
    bgot = gpu.query_rays(ctx, cam, brays)
    brows = _same_rows(bgot, bref)
    hit_frac = float(np.mean(fused["kind"] == 1))
    print(f"[configs] c3: {n} rays in {len(sizes)} batches, fused {fused_ms:.2f} ms == per-batch launches; "
          f"oracle sample {len(sample)} rays: {rows.size} differ; boundary rays {len(brays)}: {brows.size} differ; "
          f"hit fraction {hit_frac:.3f}", flush=True)
    assert rows.size == 0 and brows.size == 0
    assert 0.3 < hit_frac < 0.95


# ------------------------------------------------------------ config 4
EDIT_CHUNK_X0 = 24   # edits only in chunks x >= 24 (world x >= 153.6 m)
RAY_X_MAX_M = 76.8   # served ray batches start at x < 76.8 m and end within 51.2 m: they never reach an edited chunk


def test_c4_serving_poisson_with_edits(gpu, world_32x4x32):
    O = _oracle()
    dims, hw, scenes = world_32x4x32   # the server copies what it needs; hw itself stays the unedited world
    sc = scenes.c4(streams=8, frames=400)
    assert tuple(sc.grid_dims) == dims and sc.params.seed == hw.params.seed
    clip, heights = scenes.scene_clipmap(sc, 0)
    ctx = gpu.Context(node_capacity=int(1.5 * len(hw.nodes)), chunk_capacity=hw.chunk_count,
                      grid_capacity=len(hw.grid), clip_levels=5, clip_res=256)
    ctx.upload_world(hw)
    ctx.write_clipmap(clip, heights)
    qcam = _scene_cam(gpu, sc.grid_origin, sc.grid_dims, hw.chunk_count)
    srv = gpu.Server(ctx, hw, memory_budget=64 << 30, max_queued_frames=1, latency_target_ms=100.0)
    rng = np.random.default_rng(5)
    pool = scenes.query_rays(rng, 2_000_000, sc.grid_origin, sc.grid_dims, tmax_m=51.2)
    pool[:, 0] = rng.uniform(0.0, RAY_X_MAX_M, len(pool)).astype(np.float32)
    n_streams = len(sc.stream_poses)
    res = [scenes.VIEW_SIZES[i] for i in rng.integers(0, 4, n_streams)]   # up to 1920x1080
    rate, seconds, edit_rate = 400.0, 1.5, 1000.0
    t_arr = np.cumsum(rng.exponential(1.0 / rate, int(rate * seconds * 1.5)))
    t_arr = t_arr[t_arr < seconds]
    kinds = rng.random(len(t_arr)) < 0.5
    edits_log = []
    stop = threading.Event()
    ext_z = sc.grid_dims[2] * 64

    def editor():
        erng = np.random.default_rng(6)
        period, per = 0.02, int(edit_rate * 0.02)
        while not stop.is_set():
            x = erng.integers(EDIT_CHUNK_X0 * 64, sc.grid_dims[0] * 64, per)
            z = erng.integers(0, ext_z, per)
            y = np.array([gpu.ground_height(sc.params, int(a), int(b)) for a, b in zip(x, z)]) + erng.integers(-3, 4, per)
            xyz, mats = np.stack([x, y, z], 1).astype(np.int32), erng.integers(0, 6, per).astype(np.uint8)
            edits_log.append((xyz, mats))
            srv.edit(xyz, mats)
            stop.wait(period)

    th = threading.Thread(target=editor, daemon=True)
    frame_of = [0] * n_streams
    ray_reqs = {}
    results = []
    t0 = time.perf_counter()
    th.start()
    for i, ta in enumerate(t_arr):
        dt = t0 + ta - time.perf_counter()
        if dt > 0:
            time.sleep(dt)
        if kinds[i]:
            sid = int(rng.integers(0, n_streams))
            pose, t_s = sc.stream_poses[sid][frame_of[sid] % len(sc.stream_poses[sid])]
            w_, h_ = res[sid]
            cam = gpu.camera_build(pose, (w_, h_), (w_, h_), frame_of[sid], time_s=t_s, grid_origin=sc.grid_origin,
                                   grid_dims=sc.grid_dims, chunk_count=hw.chunk_count, jitter=True)
            if srv.submit_view(sid, cam, w_, h_, tag=i):
                frame_of[sid] += 1
        else:
            k = int(np.exp(rng.uniform(np.log(1e3), np.log(1e6))))
            off = int(rng.integers(0, len(pool) - k))
            out = np.empty(k, gpu.HIT_DTYPE)
            if srv.submit_rays(qcam, pool[off:off + k], tag=i, out=out):
                ray_reqs[i] = (off, k, out)
        while (r := srv.poll()) is not None:
            results.append(r)
    stop.set()
    th.join()
    srv.drain()
    while (r := srv.poll()) is not None:
        results.append(r)
    st = srv.stats()
    assert all(r.status == 0 for r in results), [r.status for r in results if r.status][:5]
    assert st["completed"] == st["admitted"] == len(results) and len(ray_reqs) > 50
    # every served ray batch against the oracle on the unedited world (1 % of each batch, at least 64 rays)
    w0 = O.World.from_host(hw, *worlds.empty_clipmap(gpu))
    srng = np.random.default_rng(9)
    idx, got = [], []
    for off, k, out in ray_reqs.values():
        m = max(64, k // 100)
        s = np.sort(srng.choice(k, m, replace=False))
        idx.append(off + s)
        got.append(out[s])
    idx, got = np.concatenate(idx), np.concatenate(got)
    ref = O.query_rays(w0, qcam, pool[idx], gpu.HIT_DTYPE)
    rows = _same_rows(got, ref)
    print(f"[configs] c4 concurrent phase: {len(t_arr)} arrivals, {st['admitted']} admitted, {st['rejected']} rejected, "
          f"{st['launches']} launches, view p50/p99 {st['view_p50_ms']:.1f}/{st['view_p99_ms']:.1f} ms, "
          f"{sum(len(m) for _, m in edits_log)} edits ({st['edits_applied']} chunk rebuilds applied so far); "
          f"{len(ray_reqs)} ray batches, {len(idx)} sampled rays vs oracle: {rows.size} differ", flush=True)
    assert rows.size == 0

    # checkpoint: every edit applied, then a served frame and ray batch against the oracle on a rebuilt world
    srv.wait_edits()
    all_xyz = np.concatenate([e[0] for e in edits_log])
    all_mats = np.concatenate([e[1] for e in edits_log])
    ref_world = gpu.HostWorld(sc.params, sc.grid_origin, sc.grid_dims)
    ref_world.edit(all_xyz, all_mats)
    vres = (640, 360)
    pose = gpu.make_pose((170.0, scenes.surface_m(sc.params, 170.0, 100.0) + 4.0, 100.0), 0.4, math.radians(-20.0), 70.0)
    vcam = gpu.camera_build(pose, vres, vres, 0, grid_origin=sc.grid_origin, grid_dims=sc.grid_dims,
                            chunk_count=hw.chunk_count)
    vout = np.zeros((vres[1], vres[0], 4), np.float32)
    assert srv.submit_view(999, vcam, *vres, tag=10_000, out=vout)
    all_rays = scenes.query_rays(np.random.default_rng(12), 200_000, sc.grid_origin, sc.grid_dims, tmax_m=51.2)
    ahits = np.empty(len(all_rays), gpu.HIT_DTYPE)
    assert srv.submit_rays(qcam, all_rays, tag=10_001, out=ahits)
    srv.drain()
    res2 = []
    while (r := srv.poll()) is not None:
        res2.append(r)
    assert len(res2) == 2 and all(r.status == 0 for r in res2)
    srv.close()
    w1 = O.World.from_host(ref_world, clip, heights)
    ref_hits = O.query_rays(w1, qcam, all_rays, gpu.HIT_DTYPE)
    arows = _same_rows(ahits, ref_hits)
    fr = O.Frame(*vres)
    O.frame(w1, vcam, fr)
    verr = float(np.max(np.abs(vout - fr.output)))
    edited_region = (all_rays[:, 0] >= EDIT_CHUNK_X0 * 6.4 - 51.2)
    print(f"[configs] c4 checkpoint: {len(all_xyz)} edits applied; {len(all_rays)} rays over the whole world "
          f"({int(edited_region.sum())} able to reach edited chunks): {arows.size} differ from the oracle on the "
          f"rebuilt world; served 640x360 frame in the edited region: output max abs {verr:.3g}", flush=True)
    assert arows.size == 0 and verr <= COLOR_ABS_TOL
    ctx.close()
