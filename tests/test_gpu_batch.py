"""Batched rendering (vx_batch_*): many views of mixed resolutions rendered by
fused work-stealing launches give, frame after frame, exactly the outputs of
rendering each view on its own with vx_render (which the parity tests tie to
the oracle).  Covers hipGraph replay and direct launches, several autotuner
configurations, a view rendered on its own between batch frames (history
ping-pong resynchronised), and tuning runs that must leave history as it was."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import host_world
from parity import make_context

pytestmark = pytest.mark.gpu

# (present, render): multiples and non-multiples of 8 and 16, one half-resolution view
SIZES = [((640, 360), (640, 360)), ((320, 180), (320, 180)), ((316, 180), (316, 180)), ((200, 120), (100, 60)),
         ((1280, 720), (1280, 720))]
OUTS = ("OUT_OUTPUT", "OUT_COLOR", "OUT_DEPTH", "OUT_GODRAY", "OUT_LOCAL_HIST", "OUT_PRIMARY_HIST", "OUT_PRE_TAA")


def _cams(gpu, sc, hw, frame):
    cams = []
    for i, (pwh, rwh) in enumerate(SIZES):
        pose, t_s = sc.poses[(frame + 23 * i) % len(sc.poses)]
        cams.append(gpu.camera_build(pose, rwh, pwh, frame, time_s=t_s, wind_strength=2.0, grid_origin=sc.grid_origin,
                                     grid_dims=sc.grid_dims, chunk_count=hw.chunk_count, jitter=True))
    return cams


def _same(a, b):
    return np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_batch_equals_single_views(gpu):
    from voxhip import scenes

    sc, hw = host_world("c1")
    clip, heights = scenes.scene_clipmap(sc, 0)
    ctx = make_context(gpu, hw)
    ctx.write_clipmap(clip, heights)
    single = [gpu.View(ctx, p, r) for p, r in SIZES]
    sets = {}
    for name, cfg in (("graph", dict(waves_per_simd=6, grab=1, use_graph=True)),
                      ("direct", dict(waves_per_simd=4, grab=2, use_graph=False)),
                      ("wide", dict(waves_per_simd=12, grab=4, use_graph=True))):
        views = [gpu.View(ctx, p, r) for p, r in SIZES]
        batch = gpu.Batch(ctx, views)
        batch.set_config(**cfg)
        sets[name] = (views, batch)
    for frame in range(4):
        cams = _cams(gpu, sc, hw, frame)
        for v, c in zip(single, cams):
            v.write_camera(c)
            v.render(gpu.PASS_FRAME)
        for name, (views, batch) in sets.items():
            for v, c in zip(views, cams):
                v.write_camera(c)
            if frame == 2 and name == "graph":
                batch.tune(gpu.PASS_FRAME, iters=1)  # timed runs without the flip: history unchanged
            batch.render(gpu.PASS_FRAME)
            assert batch.timing() > 0.0
        if frame == 1:
            # view 0 renders one extra frame on its own, in every set
            for vs in [single] + [s[0] for s in sets.values()]:
                vs[0].write_camera(cams[0])
                vs[0].render(gpu.PASS_FRAME)
        for name, (views, _) in sets.items():
            for i, (a, b) in enumerate(zip(single, views)):
                for o in OUTS:
                    which = getattr(gpu, o)
                    assert _same(a.read(which), b.read(which)), (name, frame, i, o)
        print(f"[batch] frame {frame}: all {len(SIZES)} views identical in {len(sets)} batch configurations", flush=True)
