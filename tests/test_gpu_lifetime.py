"""Object lifetimes and error paths at the C ABI:
  * vx_destroy refuses (VX_E_BUSY) while a view, batch, server or chunk builder
    of the context is alive, and succeeds once they are gone;
  * a server whose rebuilt chunk cannot be placed in the node arena reports
    VX_E_CAPACITY from vx_server_wait_edits instead of waiting for ever;
  * ray batches with different scene cameras in one server launch are each
    traced against their own scene."""
from __future__ import annotations

import threading

import numpy as np
import pytest

from conftest import host_world

pytestmark = pytest.mark.gpu

VX_E_BUSY = -6
VX_E_CAPACITY = -2


def _world_ctx(gpu, hw, node_capacity=None):
    from voxhip import scenes

    sc, _ = host_world("c0")
    clip, heights = scenes.scene_clipmap(sc)
    ctx = gpu.Context(node_capacity=node_capacity or 2 * len(hw.nodes), chunk_capacity=hw.chunk_count,
                      grid_capacity=len(hw.grid))
    ctx.upload_world(hw)
    ctx.write_clipmap(clip, heights)
    return ctx


def test_destroy_refused_while_children_live(gpu):
    sc, hw = host_world("c0")
    ctx = _world_ctx(gpu, hw)
    view = gpu.View(ctx, (64, 32))
    batch = gpu.Batch(ctx, [view])
    builder = gpu.ChunkBuilder(ctx, 2)
    srv = gpu.Server(ctx, hw, memory_budget=1 << 30)
    # Through the C ABI: refused while any child is alive; the context stays usable.
    for child in (srv, builder, batch, view):
        assert gpu.LIB.vx_destroy(ctx._h) == VX_E_BUSY
        assert b"still alive" in gpu.LIB.vx_last_error()
        child.close()
    view2 = gpu.View(ctx, (16, 16))  # the refused context still works
    view2.close()
    assert gpu.LIB.vx_destroy(ctx._h) == 0
    ctx._h = None


def test_destroy_refused_from_python_close(gpu):
    sc, hw = host_world("c0")
    ctx = _world_ctx(gpu, hw)
    view = gpu.View(ctx, (32, 32))
    with pytest.raises(gpu.VoxhipError, match=r"\(-6\)"):
        ctx.close()
    view.close()
    ctx.close()
    assert ctx._h is None


def test_wait_edits_reports_full_node_arena(gpu):
    """The node arena holds exactly the uploaded world, so a rebuilt chunk has
    no fresh range to go to (the old one is released only after the new one is
    written).  wait_edits must return VX_E_CAPACITY, not block."""
    from voxhip import scenes

    sc, _ = host_world("c0")
    hw = gpu.HostWorld(sc.params, sc.grid_origin, sc.grid_dims)
    ctx = _world_ctx(gpu, hw, node_capacity=len(hw.nodes))
    srv = gpu.Server(ctx, hw, memory_budget=1 << 30, edit_threads=2)
    x, z = 100, 100
    srv.edit(np.array([(x, gpu.ground_height(sc.params, x, z) + 2, z)], np.int32), np.array([6], np.uint8))
    out = {}

    def wait():
        try:
            srv.wait_edits()
            out["rc"] = 0
        except gpu.VoxhipError as e:
            out["err"] = str(e)

    th = threading.Thread(target=wait)
    th.start()
    th.join(timeout=60)
    assert not th.is_alive(), "vx_server_wait_edits did not return"
    assert "err" in out and f"({VX_E_CAPACITY})" in out["err"], out
    srv.close()
    ctx.close()


def test_ray_batches_with_different_scene_cameras(gpu):
    """Two ray batches queued together, the second with a scene camera whose
    grid origin is one chunk over: each must equal a direct query with its own
    camera (one fused launch per scene)."""
    from voxhip import scenes

    sc, hw = host_world("c0")
    ctx = _world_ctx(gpu, hw)

    def scene_cam(origin):
        return gpu.camera_build(gpu.make_pose((1.0, 1.0, 1.0), 0.0, 0.0), (64, 64), (64, 64), 0,
                                grid_origin=origin, grid_dims=sc.grid_dims, chunk_count=hw.chunk_count)

    cam_a = scene_cam(sc.grid_origin)
    cam_b = scene_cam((sc.grid_origin[0] + 1, sc.grid_origin[1], sc.grid_origin[2]))
    rays = scenes.query_rays(np.random.default_rng(11), 6000, sc.grid_origin, sc.grid_dims, tmax_m=51.2)
    direct_a = gpu.query_rays(ctx, cam_a, rays[:3000])
    direct_b = gpu.query_rays(ctx, cam_b, rays[3000:])
    assert not np.array_equal(direct_a.view(np.uint8), gpu.query_rays(ctx, cam_b, rays[:3000]).view(np.uint8))
    srv = gpu.Server(ctx, hw, memory_budget=1 << 30)
    hits = [np.zeros(3000, gpu.HIT_DTYPE) for _ in range(3)]
    assert srv.submit_rays(cam_a, rays[:1500], tag=1, out=hits[0][:1500])
    assert srv.submit_rays(cam_a, rays[1500:3000], tag=2, out=hits[0][1500:])
    assert srv.submit_rays(cam_b, rays[3000:], tag=3, out=hits[1])
    assert srv.submit_rays(cam_a, rays[:3000], tag=4, out=hits[2])
    srv.drain()
    res = []
    while (r := srv.poll()) is not None:
        res.append(r)
    assert len(res) == 4 and all(r.status == 0 for r in res)
    srv.close()
    assert np.array_equal(hits[0].view(np.uint8), direct_a.view(np.uint8))
    assert np.array_equal(hits[1].view(np.uint8), direct_b.view(np.uint8))
    assert np.array_equal(hits[2].view(np.uint8), direct_a.view(np.uint8))
    ctx.close()
