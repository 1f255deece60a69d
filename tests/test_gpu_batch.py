"""Batched rendering (vx_batch_*): many views of mixed resolutions rendered by
fused work-stealing launches give, frame after frame, exactly the outputs of
rendering each view on its own with vx_render (which the parity tests tie to
the oracle).  Covers hipGraph replay and direct launches, several autotuner
configurations including the 16x4 and 32x2 execution tiles (each pixel with
its 8x8 tile's candidate list), a view rendered on its own between batch
frames (history ping-pong resynchronised), and tuning runs that must leave
history as it was."""
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
                      ("wide", dict(waves_per_simd=12, grab=4, use_graph=True)),
                      ("16x4", dict(waves_per_simd=4, grab=1, use_graph=True, tile_shape=1)),
                      ("32x2", dict(waves_per_simd=8, grab=2, use_graph=False, tile_shape=2))):
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


def _random_rays(rng, n, grid_dims, tmax=51.2):
    """Origins uniform in the world AABB, directions uniform on the sphere (the ray-query workload's shape)."""
    ext = np.array(grid_dims, np.float64) * 6.4
    o = rng.uniform(0.0, 1.0, (n, 3)) * ext
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3], r[:, 3], r[:, 4:7] = o, tmax, d
    return r


def test_batch_with_fused_rays(gpu):
    """Ray batches fused into a batch launch (vx_batch_set_rays): the hits
    equal vx_query_rays on the same rays byte for byte (and the oracle on a
    sample and on the boundary rays), the views' outputs are those of the
    same batch without rays, replacing the rays between frames takes effect,
    and n = 0 detaches them; with a graph and with direct launches."""
    import worlds
    from oracle import oracle as O
    from voxhip import scenes

    sc, hw = host_world("c1")
    clip, heights = scenes.scene_clipmap(sc, 0)
    ctx = make_context(gpu, hw)
    ctx.write_clipmap(clip, heights)
    qcam = gpu.camera_build(gpu.make_pose((1.0, 1.0, 1.0), 0.0, 0.0), (64, 64), (64, 64), 0, grid_origin=sc.grid_origin,
                            grid_dims=sc.grid_dims, chunk_count=hw.chunk_count)
    rng = np.random.default_rng(17)
    # the first set is large enough to be sorted before it is traced (2^20 rays or more), the second is not
    n_random = 1_100_000
    ray_sets = [np.concatenate([_random_rays(rng, n_random, sc.grid_dims), worlds.boundary_rays()]),
                _random_rays(rng, 70_001, sc.grid_dims)]
    w = O.World.from_host(hw, clip, heights)
    for cfg in (dict(waves_per_simd=4, grab=1, use_graph=True), dict(waves_per_simd=6, grab=2, use_graph=False)):
        plain_views = [gpu.View(ctx, p, r) for p, r in SIZES]
        fused_views = [gpu.View(ctx, p, r) for p, r in SIZES]
        plain, fused = gpu.Batch(ctx, plain_views), gpu.Batch(ctx, fused_views)
        plain.set_config(**cfg)
        fused.set_config(**cfg)
        for frame in range(3):
            rays = ray_sets[frame % 2]
            n = len(rays)
            d_rays = gpu.DeviceArray.from_numpy(rays)
            d_hits = gpu.DeviceArray.from_numpy(np.zeros(n, gpu.HIT_DTYPE))
            d_ref = gpu.DeviceArray(n * gpu.HIT_DTYPE.itemsize)
            fused.set_rays(qcam, d_rays, d_hits, n if frame < 2 else 0)
            cams = _cams(gpu, sc, hw, frame)
            for vs in (plain_views, fused_views):
                for v, c in zip(vs, cams):
                    v.write_camera(c)
            plain.render(gpu.PASS_FRAME)
            fused.render(gpu.PASS_FRAME)
            gpu.query_rays_device(ctx, qcam, d_rays, d_ref, n)
            ctx.sync()
            for i, (a, b) in enumerate(zip(plain_views, fused_views)):
                for o in OUTS:
                    which = getattr(gpu, o)
                    assert _same(a.read(which), b.read(which)), (cfg, frame, i, o)
            ref = d_ref.download(np.empty(n, gpu.HIT_DTYPE))
            got = d_hits.download(np.empty(n, gpu.HIT_DTYPE))
            if frame < 2:
                assert np.array_equal(got.view(np.uint8), ref.view(np.uint8)), (cfg, frame)
                sample = np.sort(np.random.default_rng(frame).choice(n, 4000, replace=False))
                if frame == 0:
                    sample = np.union1d(sample, np.arange(n_random, n))  # the boundary rays
                oref = O.query_rays(w, qcam, rays[sample], gpu.HIT_DTYPE)
                assert np.array_equal(got[sample].view(np.uint8), oref.view(np.uint8)), (cfg, frame)
                hit_frac = float(np.mean(got["kind"] == 1))
                assert 0.2 < hit_frac < 0.98
            else:
                assert not got.view(np.uint8).any()   # detached: the launch traced no rays
            for d in (d_rays, d_hits, d_ref):
                d.close()
            print(f"[batch] fused rays {cfg}: frame {frame}: {n} rays, views identical to the batch without rays",
                  flush=True)
        fused.set_rays(None, None, None, 0)
        plain.close()
        fused.close()
