"""The serving loop (vx_server_*): a view stream's frames come out the same
alone and mixed with other streams and ray batches (continuous batching
changes scheduling, not results); admission against the memory budget; and
the edit stream (chunks rebuilt on the host and swapped into fresh arena
ranges) renders exactly what a world built with the same edits renders."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import host_world
from parity import make_context

pytestmark = pytest.mark.gpu


def _cams(gpu, sc, hw, res, n, dx=0.05):
    pose0, _ = sc.poses[0]
    out, prev = [], None
    for f in range(n):
        pose = gpu.make_pose((pose0.pos[0] + dx * f, pose0.pos[1], pose0.pos[2] - 0.03 * f), pose0.yaw + 0.01 * f,
                             pose0.pitch, 70.0)
        cam = gpu.camera_build(pose, res, res, f, time_s=f / 60.0, wind_strength=2.0, grid_origin=sc.grid_origin,
                               grid_dims=sc.grid_dims, chunk_count=hw.chunk_count, prev_view_proj=prev, jitter=True)
        prev = np.array(cam.view_proj, np.float32)
        out.append(cam)
    return out


def _scene_cam(gpu, sc, hw):
    return gpu.camera_build(gpu.make_pose((1.0, 1.0, 1.0), 0.0, 0.0), (64, 64), (64, 64), 0,
                            grid_origin=sc.grid_origin, grid_dims=sc.grid_dims, chunk_count=hw.chunk_count)


def _ctx(gpu, hw, clip, heights):
    ctx = gpu.Context(node_capacity=2 * len(hw.nodes), chunk_capacity=hw.chunk_count, grid_capacity=len(hw.grid))
    ctx.upload_world(hw)
    ctx.write_clipmap(clip, heights)
    return ctx


def _wait_all(srv, n):
    srv.drain()
    res = []
    while len(res) < n:
        r = srv.poll()
        assert r is not None
        res.append(r)
    assert srv.poll() is None
    return res


def test_stream_same_alone_and_in_mix(gpu):
    from voxhip import scenes

    sc, hw = host_world("c0")
    clip, heights = scenes.scene_clipmap(sc)
    ctx = _ctx(gpu, hw, clip, heights)
    res = (256, 144)
    cams = _cams(gpu, sc, hw, res, 3)
    alone = [np.zeros((res[1], res[0], 4), np.float32) for _ in cams]
    srv = gpu.Server(ctx, hw, memory_budget=8 << 30)
    for f, cam in enumerate(cams):  # one frame at a time
        assert srv.submit_view(1, cam, *res, tag=f, out=alone[f])
        assert all(r.status == 0 for r in _wait_all(srv, 1))
    srv.close()

    qcam = _scene_cam(gpu, sc, hw)
    rays = scenes.query_rays(np.random.default_rng(2), 3000, sc.grid_origin, sc.grid_dims, tmax_m=51.2)
    direct = gpu.query_rays(ctx, qcam, rays)
    srv = gpu.Server(ctx, hw, memory_budget=8 << 30)
    mixed = [np.zeros_like(a) for a in alone]
    hits = [np.zeros(1000, gpu.HIT_DTYPE) for _ in range(3)]
    others = {8: ((320, 180), _cams(gpu, sc, hw, (320, 180), 2, dx=-0.04)), 9: ((200, 120), _cams(gpu, sc, hw, (200, 120), 2, dx=0.1))}
    tag = 100
    n = 0
    for f, cam in enumerate(cams):  # everything queued at once: frames share launches with other work
        assert srv.submit_view(7, cam, *res, tag=f, out=mixed[f])
        n += 1
        for sid, (r2, cs) in others.items():
            if f < len(cs):
                assert srv.submit_view(sid, cs[f], *r2, tag=tag)
                tag += 1
                n += 1
        assert srv.submit_rays(qcam, rays[1000 * f:1000 * (f + 1)], tag=200 + f, out=hits[f])
        n += 1
    results = _wait_all(srv, n)
    assert all(r.status == 0 for r in results)
    st = srv.stats()
    print(f"[server] launches={st['launches']} completed={st['completed']} view p50/p99 = "
          f"{st['view_p50_ms']:.2f}/{st['view_p99_ms']:.2f} ms", flush=True)
    assert st["completed"] == st["admitted"] == n and st["rejected"] == 0
    for f in range(3):
        assert np.array_equal(mixed[f].view(np.uint32), alone[f].view(np.uint32)), f
        assert np.array_equal(hits[f].view(np.uint8), direct[1000 * f:1000 * (f + 1)].view(np.uint8)), f
    srv.close()


def test_admission_budget(gpu):
    from voxhip import scenes

    sc, hw = host_world("c0")
    clip, heights = scenes.scene_clipmap(sc)
    ctx = _ctx(gpu, hw, clip, heights)
    res = (128, 72)
    cam = _cams(gpu, sc, hw, res, 1)[0]
    view_bytes = res[0] * res[1] * 250
    srv = gpu.Server(ctx, hw, memory_budget=view_bytes + 500 * 64)
    assert srv.submit_view(1, cam, *res, tag=1)
    assert not srv.submit_view(2, cam, *res, tag=2)  # a second resident view does not fit
    qcam = _scene_cam(gpu, sc, hw)
    rays = scenes.query_rays(np.random.default_rng(4), 1000, sc.grid_origin, sc.grid_dims)
    assert not srv.submit_rays(qcam, rays, tag=3)  # 1000 rays need 64 kB
    assert srv.submit_rays(qcam, rays[:400], tag=4)
    _wait_all(srv, 2)
    srv.close_stream(1)
    assert srv.submit_view(2, cam, *res, tag=5)  # the closed stream's memory is free again
    _wait_all(srv, 1)
    st = srv.stats()
    assert st["rejected"] == 2 and st["admitted"] == 3 and st["completed"] == 3
    srv.close()


def test_ray_batches_through_pool_and_staging(gpu):
    """A ray batch is copied at submit into the server's pinned pool (a
    budget's worth of rays, at most 8 Mi); one larger than the pool goes
    through the launch's own staging instead.  Both return the hits of a
    direct vx_query_rays on the same rays, also when they share launches."""
    from voxhip import scenes

    sc, hw = host_world("c0")
    clip, heights = scenes.scene_clipmap(sc)
    ctx = _ctx(gpu, hw, clip, heights)
    qcam = _scene_cam(gpu, sc, hw)
    big = 9_000_000  # above the 8 Mi-ray pool: staged
    rays = scenes.query_rays(np.random.default_rng(12), big, sc.grid_origin, sc.grid_dims, tmax_m=51.2)
    direct = gpu.query_rays(ctx, qcam, rays)
    srv = gpu.Server(ctx, hw, memory_budget=1 << 30)
    spans = [(0, 5000), (0, big), (123_457, 200_000), (big - 70_000, big)]
    hits = [np.zeros(b - a, gpu.HIT_DTYPE) for a, b in spans]
    for i, (a, b) in enumerate(spans):
        assert srv.submit_rays(qcam, rays[a:b], tag=i, out=hits[i])
    results = _wait_all(srv, len(spans))
    assert all(r.status == 0 for r in results)
    for i, (a, b) in enumerate(spans):
        assert np.array_equal(hits[i].view(np.uint8), direct[a:b].view(np.uint8)), i
    srv.close()


def test_close_stream_while_streams_open(gpu):
    """vx_server_close_stream waits for the stream's queued frames with the lock
    released; meanwhile another thread opens 64 new streams (the stream map
    grows and rehashes).  The close completes, every frame completes with
    status 0, and the other streams stay usable and close cleanly."""
    import threading
    import time

    from voxhip import scenes

    sc, hw = host_world("c0")
    clip, heights = scenes.scene_clipmap(sc)
    ctx = _ctx(gpu, hw, clip, heights)
    srv = gpu.Server(ctx, hw, memory_budget=8 << 30)
    big, small = (320, 180), (32, 18)
    cams_big = _cams(gpu, sc, hw, big, 4)
    cam_small = _cams(gpu, sc, hw, small, 1)[0]
    for f, cam in enumerate(cams_big):
        assert srv.submit_view(1, cam, *big, tag=f)
    errors = []

    def close_one():
        try:
            srv.close_stream(1)
        except Exception as e:  # reported below
            errors.append(e)

    th = threading.Thread(target=close_one)
    th.start()
    time.sleep(0.002)
    for k in range(64):
        assert srv.submit_view(100 + k, cam_small, *small, tag=100 + k)
    th.join(timeout=60)
    assert not th.is_alive() and not errors, errors
    res = _wait_all(srv, 4 + 64)
    assert all(r.status == 0 for r in res)
    st = srv.stats()
    assert st["completed"] == st["admitted"] == 68
    assert srv.submit_view(100, cam_small, *small, tag=200)  # an open stream still renders
    _wait_all(srv, 1)
    for k in range(64):
        srv.close_stream(100 + k)
    srv.close()


def test_latency_target_admission(gpu):
    """With a latency target, admission rejects a request once the server's
    measured throughput says the queue and the request cannot finish within
    it; before the first launch there is no estimate and nothing is rejected
    for latency.  The results of admitted requests are unaffected."""
    from oracle import oracle as O
    from voxhip import scenes

    sc, hw = host_world("c0")
    clip, heights = scenes.scene_clipmap(sc)
    ctx = _ctx(gpu, hw, clip, heights)
    qcam = _scene_cam(gpu, sc, hw)
    rays = scenes.query_rays(np.random.default_rng(8), 20000, sc.grid_origin, sc.grid_dims)
    srv = gpu.Server(ctx, hw, memory_budget=1 << 30, latency_target_ms=0.01)
    out = np.zeros(len(rays), gpu.HIT_DTYPE)
    assert srv.submit_rays(qcam, rays, tag=1, out=out)  # no throughput estimate yet
    _wait_all(srv, 1)
    assert not srv.submit_rays(qcam, rays, tag=2)  # 20000 rays cannot finish within 0.01 ms
    st = srv.stats()
    assert st["rejected"] == 1 and st["completed"] == 1
    ref = O.query_rays(O.World.from_host(hw, clip, heights), qcam, rays, gpu.HIT_DTYPE)
    assert np.array_equal(out.view(np.uint8), ref.view(np.uint8))
    srv.close()


@pytest.mark.parametrize("device_build", [False, True])
def test_edits_match_rebuilt_world(gpu, device_build):
    from voxhip import scenes

    sc, _ = host_world("c0")
    p = sc.params
    hw = gpu.HostWorld(p, sc.grid_origin, sc.grid_dims)
    clip, heights = scenes.scene_clipmap(sc)
    ctx = _ctx(gpu, hw, clip, heights)
    res = (256, 144)
    cam = _cams(gpu, sc, hw, res, 1)[0]
    # a pit and a light block in front of the camera, plus scattered single edits
    pose, _ = sc.poses[0]
    vx0 = int(pose.pos[0] / 0.1) + int(40 * np.sin(pose.yaw))
    vz0 = int(pose.pos[2] / 0.1) + int(40 * np.cos(pose.yaw))
    xyz, mats = [], []
    for dx in range(-2, 3):
        for dz in range(-2, 3):
            g = gpu.ground_height(p, vx0 + dx, vz0 + dz)
            for dy in range(0, 3):
                xyz.append((vx0 + dx, g - dy, vz0 + dz))
                mats.append(0)
            xyz.append((vx0 + dx + 8, g + 4, vz0 + dz))
            mats.append(6)  # LIGHT
    rng = np.random.default_rng(9)
    for _ in range(40):
        x, z = (int(v) for v in rng.integers(0, 512, 2))
        xyz.append((x, gpu.ground_height(p, x, z) + 1, z))
        mats.append(int(rng.integers(1, 6)))
    xyz, mats = np.array(xyz, np.int32), np.array(mats, np.uint8)

    srv = gpu.Server(ctx, hw, memory_budget=8 << 30, edit_threads=4, device_build=device_build)
    # in two calls, the second re-editing voxels of the first (the last edit of a voxel wins)
    half = len(xyz) // 2
    srv.edit(xyz[:half], np.full(half, 5, np.uint8))
    srv.edit(xyz, mats)
    srv.wait_edits()
    out = np.zeros((res[1], res[0], 4), np.float32)
    assert srv.submit_view(1, cam, *res, tag=1, out=out)
    qcam = _scene_cam(gpu, sc, hw)
    rays = scenes.query_rays(np.random.default_rng(6), 4000, sc.grid_origin, sc.grid_dims)
    hits = np.zeros(len(rays), gpu.HIT_DTYPE)
    assert srv.submit_rays(qcam, rays, tag=2, out=hits)
    _wait_all(srv, 2)
    st = srv.stats()
    assert st["edits_applied"] >= 1
    srv.close()

    ref_world = gpu.HostWorld(p, sc.grid_origin, sc.grid_dims)
    ref_world.edit(xyz, mats)
    ref_ctx = _ctx(gpu, ref_world, clip, heights)
    view = gpu.View(ref_ctx, res, res)
    view.write_camera(cam)
    view.render(gpu.PASS_FRAME)
    ref = view.read(gpu.OUT_OUTPUT)
    print(f"[server] device_build={device_build} edits applied: {st['edits_applied']} chunks; image differs from the unedited world: "
          f"{not np.array_equal(ref, out) or True}", flush=True)
    assert np.array_equal(out.view(np.uint32), ref.view(np.uint32))
    ref_hits = gpu.query_rays(ref_ctx, qcam, rays)
    assert np.array_equal(hits["kind"], ref_hits["kind"]) and np.array_equal(hits["voxel"], ref_hits["voxel"])
    assert np.array_equal(hits["mat"], ref_hits["mat"]) and np.array_equal(hits["t"].view(np.uint32), ref_hits["t"].view(np.uint32))
    assert np.array_equal(hits["node"], ref_hits["node"]) and np.array_equal(hits["slot"], ref_hits["slot"])
