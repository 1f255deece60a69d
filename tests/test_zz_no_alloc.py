"""vx_render_views_noalloc, the entry point that allocates nothing: its
scratch (view table, ping words, tile cursors) is a static device array, so
callers that cannot allocate can render several views in one fused launch.
Two calls in a row, the second with fewer and smaller views, must equal a
vx_batch over the same views (vx_batch_render) frame for frame, bit for bit;
more views than the scratch holds are rejected."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from conftest import host_world
from parity import make_context

pytestmark = pytest.mark.gpu

OUTS = ("OUT_OUTPUT", "OUT_COLOR", "OUT_DEPTH", "OUT_GODRAY", "OUT_LOCAL_HIST", "OUT_PRIMARY_HIST", "OUT_PRE_TAA")


def _cams(gpu, sc, hw, sizes, frame):
    out = []
    for i, (pwh, rwh) in enumerate(sizes):
        pose, t_s = sc.poses[(frame + 31 * i) % len(sc.poses)]
        out.append(gpu.camera_build(pose, rwh, pwh, frame, time_s=t_s, wind_strength=2.0, grid_origin=sc.grid_origin,
                                    grid_dims=sc.grid_dims, chunk_count=hw.chunk_count, jitter=True))
    return out


def _noalloc(gpu, ctx, views, mask):
    arr = (C.c_void_p * max(1, len(views)))(*[v._h.value for v in views])
    return gpu.LIB.vx_render_views_noalloc(ctx._h, arr, len(views), mask, None)


def test_render_views_noalloc_equals_batch(gpu):
    from voxhip import scenes

    sc, hw = host_world("c1")
    clip, heights = scenes.scene_clipmap(sc, 0)
    ctx = make_context(gpu, hw)
    ctx.write_clipmap(clip, heights)
    calls = [
        [((640, 360), (640, 360)), ((320, 180), (320, 180)), ((200, 120), (100, 60)), ((316, 180), (316, 180))],
        [((160, 90), (160, 90)), ((96, 64), (96, 64))],   # the second call: fewer, smaller views
    ]
    for k, sizes in enumerate(calls):
        free = [gpu.View(ctx, p, r) for p, r in sizes]    # rendered by the no-alloc entry point
        ref = [gpu.View(ctx, p, r) for p, r in sizes]     # rendered by a vx_batch
        batch = gpu.Batch(ctx, ref)
        for frame in range(3):
            cams = _cams(gpu, sc, hw, sizes, frame + 10 * k)
            for a, b, c in zip(free, ref, cams):
                a.write_camera(c)
                b.write_camera(c)
            assert _noalloc(gpu, ctx, free, gpu.PASS_FRAME) == 0, gpu.LIB.vx_last_error()
            batch.render(gpu.PASS_FRAME)
            for i, (a, b) in enumerate(zip(free, ref)):
                for name in OUTS:
                    which = getattr(gpu, name)
                    x, y = a.read(which), b.read(which)
                    assert np.array_equal(x.view(np.uint32), y.view(np.uint32)), (k, frame, i, name)
        print(f"[no-alloc] call {k}: {len(sizes)} views, 3 frames equal to vx_batch_render", flush=True)
        batch.close()
        for v in free + ref:
            v.close()


def test_render_views_noalloc_rejects_too_many_views(gpu):
    sc, hw = host_world("c0")
    ctx = make_context(gpu, hw)
    views = [gpu.View(ctx, (16, 16)) for _ in range(gpu.NOALLOC_MAX_VIEWS + 1)]
    assert _noalloc(gpu, ctx, views, gpu.PASS_FRAME) == gpu.VX_E_CAPACITY
    assert b"at most 64" in gpu.LIB.vx_last_error()
    assert _noalloc(gpu, ctx, views[:gpu.NOALLOC_MAX_VIEWS], gpu.PASS_FRAME) == 0   # exactly the maximum is fine
    ctx.sync()
    for v in views:
        v.close()
