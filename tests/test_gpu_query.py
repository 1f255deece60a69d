"""Ray queries on the GPU (vx_query_rays): the persistent lane-refill kernel
against the oracle's query (itself pinned to a dense brute-force walk in
test_query_oracle.py), bit-exact hit records, under several kernel
configurations; plus the edge cases of the interface."""
from __future__ import annotations

import numpy as np
import pytest

import worlds
from conftest import host_world
from parity import make_context

pytestmark = pytest.mark.gpu


def _scene_cam(gpu, sc, hw):
    return gpu.camera_build(gpu.make_pose((1.0, 1.0, 1.0), 0.0, 0.0, 70.0), (64, 64), (64, 64), 0,
                            grid_origin=sc.grid_origin, grid_dims=sc.grid_dims, chunk_count=hw.chunk_count)


def test_query_matches_oracle(gpu):
    from oracle import oracle as O
    from voxhip import scenes

    sc, hw = host_world("c1")
    ctx = make_context(gpu, hw)
    cam = _scene_cam(gpu, sc, hw)
    rays = scenes.query_rays(np.random.default_rng(11), 100_000, sc.grid_origin, sc.grid_dims, tmax_m=51.2)
    # edge cases: axis-aligned directions, zero range, origin far outside the grid
    rays[0, 4:7] = (1.0, 0.0, 0.0)
    rays[1, 4:7] = (0.0, -1.0, 0.0)
    rays[2, 3] = 0.0
    rays[3, 0:3] = (-500.0, 10.0, 20.0)
    rays[3, 4:7] = (1.0, 0.0, 0.0)
    rays[3, 3] = 1000.0
    clip, heights = worlds.empty_clipmap(gpu)
    ref = O.query_rays(O.World.from_host(hw, clip, heights), cam, rays, gpu.HIT_DTYPE)
    # sort_rays on (the default) or off: 100 k rays stay below the sort's 2^20 threshold either way
    for waves, refill, sort in ((8, 48, True), (4, 1, False), (6, 64, True)):
        gpu.query_config(ctx, waves_per_simd=waves, refill_below=refill, sort_rays=sort)
        got = gpu.query_rays(ctx, cam, rays)
        bad = np.nonzero(got.view(np.uint8).reshape(len(rays), -1).any(axis=1) !=
                         ref.view(np.uint8).reshape(len(rays), -1).any(axis=1))[0]
        same = np.array_equal(got.view(np.uint8), ref.view(np.uint8))
        print(f"[query] waves={waves} refill={refill} sort={sort}: identical={same} hits={int(np.sum(got['kind'] == 1))} "
              f"ms={gpu.query_timing(ctx):.3f}", flush=True)
        assert same, (waves, refill, sort, bad[:5])
    assert int(np.sum(ref["kind"] == 1)) > 10_000


def test_query_sorted_equals_input_order(gpu):
    """A call of 2^20 rays or more is traced sorted by direction and origin:
    its hits equal those of the same call in input order byte for byte, and
    the oracle's on a sample."""
    from oracle import oracle as O
    from voxhip import scenes

    sc, hw = host_world("c1")
    ctx = make_context(gpu, hw)
    cam = _scene_cam(gpu, sc, hw)
    rays = scenes.query_rays(np.random.default_rng(23), 1_100_000, sc.grid_origin, sc.grid_dims, tmax_m=51.2)
    got = {}
    for sort in (True, False):
        gpu.query_config(ctx, waves_per_simd=6, refill_below=16, sort_rays=sort)
        got[sort] = gpu.query_rays(ctx, cam, rays)
        print(f"[query] 1.1 M rays sort={sort}: {gpu.query_timing(ctx):.3f} ms", flush=True)
    assert np.array_equal(got[True].view(np.uint8), got[False].view(np.uint8))
    sample = np.sort(np.random.default_rng(1).choice(len(rays), 3000, replace=False))
    clip, heights = worlds.empty_clipmap(gpu)
    ref = O.query_rays(O.World.from_host(hw, clip, heights), cam, rays[sample], gpu.HIT_DTYPE)
    assert np.array_equal(got[True][sample].view(np.uint8), ref.view(np.uint8))


def test_query_boundary_rays_match_oracle(gpu):
    """Rays whose origins and directions sit on the traversal's ties (slab
    ties, exit-face ambiguity and re-descent, inward-biased chunk coordinates,
    the parallel cut): hit records bit-identical to the oracle's."""
    from oracle import oracle as O

    sc, hw = host_world("c1")
    ctx = make_context(gpu, hw)
    cam = _scene_cam(gpu, sc, hw)
    rays = worlds.boundary_rays()
    clip, heights = worlds.empty_clipmap(gpu)
    ref = O.query_rays(O.World.from_host(hw, clip, heights), cam, rays, gpu.HIT_DTYPE)
    got = gpu.query_rays(ctx, cam, rays)
    rows = np.nonzero((got.view(np.uint8).reshape(len(rays), -1) != ref.view(np.uint8).reshape(len(rays), -1))
                      .any(axis=1))[0]
    kinds = np.bincount(ref["kind"].ravel(), minlength=2).tolist()
    print(f"[query] boundary rays: {len(rays)}, kinds {kinds}, mismatched rows {rows[:8].tolist()}", flush=True)
    assert rows.size == 0, [(rays[r].tolist(), got[r], ref[r]) for r in rows[:3]]
    # both outcomes occur: rays leaving the ground hit, rays into open sky miss
    assert 0 < kinds[1] < len(rays)
    ctx.close()


def test_query_counted_loads_match_oracle(gpu):
    """vx_query_count, the counting form behind the query roofline's
    algorithmic bytes: every load counter equal to the oracle's own counts for
    the same rays, the ray and hit records counted once each, the hits equal
    to vx_query_rays', and the executed loads no more than the reference's."""
    from oracle import oracle as O
    from voxhip import scenes

    sc, hw = host_world("c1")
    ctx = make_context(gpu, hw)
    cam = _scene_cam(gpu, sc, hw)
    rays = scenes.query_rays(np.random.default_rng(5), 20_000, sc.grid_origin, sc.grid_dims, tmax_m=51.2)
    n = len(rays)
    clip, heights = worlds.empty_clipmap(gpu)
    oc = O.Counts()
    ref = O.query_rays(O.World.from_host(hw, clip, heights), cam, rays, gpu.HIT_DTYPE, counts=oc)
    d_rays = gpu.DeviceArray.from_numpy(rays)
    d_hits = gpu.DeviceArray(n * gpu.HIT_DTYPE.itemsize)
    try:
        c_ref, p_ref = gpu.query_count(ctx, cam, d_rays, d_hits, n)
        hits_ref = d_hits.download(np.empty(n, gpu.HIT_DTYPE))
        c_exe, _ = gpu.query_count(ctx, cam, d_rays, d_hits, n, executed=True)
        hits_exe = d_hits.download(np.empty(n, gpu.HIT_DTYPE))
    finally:
        d_rays.close()
        d_hits.close()
    print(f"[query] counted loads: device {c_ref}, executed {c_exe}, oracle {oc.as_dict()}", flush=True)
    for k in ("grid", "meta", "macro", "node", "rope", "colinfo", "clip"):
        assert c_ref[k] == oc.as_dict()[k], k
        assert c_exe[k] <= c_ref[k], k
    assert c_ref["tex_rd"] == 32 * n and c_ref["tex_wr"] == 32 * n
    assert sum(p_ref.values()) == gpu.algorithmic_bytes(c_ref)
    assert np.array_equal(hits_ref.view(np.uint8), ref.view(np.uint8))
    assert np.array_equal(hits_exe.view(np.uint8), ref.view(np.uint8))
    assert c_ref["node"] > 0 and c_ref["rope"] > 0


def test_query_interface_edges(gpu):
    sc, hw = host_world("c0")
    ctx = make_context(gpu, hw)
    cam = _scene_cam(gpu, sc, hw)
    d = gpu.DeviceArray(64)
    gpu.query_rays_device(ctx, cam, d, d, 0)  # n = 0: nothing to do
    ctx.sync()
    counts, parts = gpu.query_count(ctx, cam, d, d, 0)  # n = 0: no loads
    assert not any(counts.values()) and not any(parts.values())
    with pytest.raises(gpu.VoxhipError):
        gpu.query_config(ctx, waves_per_simd=0, refill_below=99)
    d.close()
