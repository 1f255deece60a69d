"""GPU: the chunk builder on the device (vx_chunk_builder_*) against the host
builder (vx_build_chunk, itself byte-exact against the numpy restatement of the
reference builder in oracle/svo_builder.py): nodes, ropes, macro and colinfo
words byte for byte, for world chunks, random volumes, collapse with orphaned
subtrees, empty / uniform / maximal chunks, in one batch at scattered arena
offsets; then a whole world built on the device equals the host-built arenas.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import host_world
from test_builder import _chunks_of_interest, _random_volume

pytestmark = pytest.mark.gpu


def _cases(vx):
    out = []
    for p, cx, cy, cz in _chunks_of_interest(vx):
        m, g = vx.fill_chunk(p, cx, cy, cz)
        out.append((f"world {cx},{cy},{cz}", m, g, cy * 64, p.dirt_depth))
    for seed in range(4):
        rng = np.random.default_rng(seed)
        m, g = _random_volume(rng, density=[0.0, 0.001, 0.02, 0.2][seed])
        out.append((f"random {seed}", m, g, 0, 4))
    # mixed solids below the stone line: full cubes collapse to STONE over interior children
    m = np.zeros((64, 64, 64), np.uint8)
    m[:32] = 3
    m[0:8, 0:8, 0:8] = 2
    m[5, 5, 5] = 1
    m[40:44, 10:20, 3:9] = 5
    out.append(("collapse over mixed solids", m.reshape(-1), np.full(4096, 100, np.int32), 0, 4))
    out.append(("empty", np.zeros(64 ** 3, np.uint8), np.zeros(4096, np.int32), 0, 4))
    out.append(("uniform dirt above ground", np.full(64 ** 3, 2, np.uint8), np.zeros(4096, np.int32), 0, 4))
    # a 3-D checkerboard: every cell interior down to the voxels, the maximal node count
    y, z, x = np.meshgrid(np.arange(64), np.arange(64), np.arange(64), indexing="ij")
    checker = (((x + y + z) & 1) + 1).astype(np.uint8).reshape(-1)
    out.append(("checkerboard", checker, np.zeros(4096, np.int32), 0, 4))
    return out


def test_device_builder_matches_host_builder(gpu):
    cases = _cases(gpu)
    refs = [gpu.build_chunk(m, g, oy, dd) for _, m, g, oy, dd in cases]
    n = len(cases)
    counts_ref = np.array([len(r.nodes) for r in refs], np.uint64)
    assert counts_ref[-1] == gpu.MAX_CHUNK_NODES
    # scattered placement: a gap before every chunk, slots in reverse order
    gaps = 37 * (np.arange(n, dtype=np.uint64) + 1)
    node_base = np.concatenate([[0], np.cumsum(counts_ref[:-1] + gaps[:-1])]).astype(np.uint64) + gaps[0]
    slots = np.arange(n)[::-1].astype(np.uint32)
    cap = int(node_base[-1] + counts_ref[-1] + 5)
    ctx = gpu.Context(node_capacity=cap, chunk_capacity=n, grid_capacity=1)
    mats = gpu.DeviceArray.from_numpy(np.concatenate([c[1] for c in cases]))
    grounds = gpu.DeviceArray.from_numpy(np.concatenate([c[2] for c in cases]).astype(np.int32))
    oys = [c[3] for c in cases]
    assert len(set(c[4] for c in cases)) == 1
    b = gpu.ChunkBuilder(ctx, max_chunks=16)
    counts = b.count(mats, grounds, oys, cases[0][4])
    assert np.array_equal(counts.astype(np.uint64), counts_ref), (counts, counts_ref)
    b.emit(node_base, slots * 16, slots * 2048)
    for i, (name, *_rest) in enumerate(cases):
        r = refs[i]
        nodes = ctx.read_region(gpu.ARENA_NODES, int(node_base[i]), len(r.nodes))
        ropes = ctx.read_region(gpu.ARENA_ROPES, int(node_base[i]), len(r.ropes))
        macro = ctx.read_region(gpu.ARENA_MACRO, int(slots[i]) * 16, 16)
        colinfo = ctx.read_region(gpu.ARENA_COLINFO, int(slots[i]) * 2048, 2048)
        assert np.array_equal(nodes.view(np.uint32), r.nodes.view(np.uint32)), name
        assert np.array_equal(ropes.view(np.uint32), r.ropes.view(np.uint32)), name
        assert np.array_equal(macro, r.macro), name
        assert np.array_equal(colinfo, r.colinfo), name
    # the gaps between chunks are untouched (zero-filled at allocation)
    for i in range(n):
        start = 0 if i == 0 else int(node_base[i - 1] + counts_ref[i - 1])
        gap = ctx.read_region(gpu.ARENA_NODES, start, int(node_base[i]) - start)
        assert not gap.view(np.uint32).any(), i
    print(f"[builder] {n} chunks, node counts {counts.tolist()}", flush=True)
    b.close()
    ctx.close()


def test_builder_argument_errors(gpu):
    ctx = gpu.Context(node_capacity=64, chunk_capacity=2, grid_capacity=1)
    b = gpu.ChunkBuilder(ctx, max_chunks=2)
    mats = gpu.DeviceArray.from_numpy(np.zeros(3 * 64 ** 3, np.uint8))
    grounds = gpu.DeviceArray.from_numpy(np.zeros(3 * 4096, np.int32))
    with pytest.raises(gpu.VoxhipError, match="no counted batch"):
        b.emit([0], [0], [0])
    with pytest.raises(gpu.VoxhipError, match="chunks"):
        b.count(mats, grounds, [0, 0, 0], 4)
    counts = b.count(mats, grounds, [0, 0], 4)
    assert counts.tolist() == [1, 1]  # empty chunks: the root alone
    with pytest.raises(gpu.VoxhipError, match="capacity"):
        b.emit([0, 64], [0, 16], [0, 2048])  # node range past the arena
    with pytest.raises(gpu.VoxhipError, match="capacity"):
        b.emit([0, 1], [0, 32], [0, 2048])  # macro base past the chunk capacity
    b.emit([0, 1], [0, 16], [0, 2048])
    nodes = ctx.read_region(gpu.ARENA_NODES, 0, 2)
    assert nodes["child_base"].tolist() == [0xFFFFFFFF] * 2 and nodes["material"].tolist() == [0, 0]
    assert nodes["key"].tolist() == [0, 0]
    b.close()
    ctx.close()


def test_world_built_on_device_equals_host_world(gpu):
    """Every slot of the seeded 8x4x8-chunk world built on the device at the
    host world's own placement: the whole arenas are byte-identical."""
    sc, hw = host_world("c0")
    meta = hw.meta
    p = sc.params
    n = len(meta)
    ctx = gpu.Context(node_capacity=len(hw.nodes), chunk_capacity=n, grid_capacity=len(hw.grid))
    b = gpu.ChunkBuilder(ctx, max_chunks=128)
    mats = gpu.DeviceArray(128 * 64 ** 3)
    grounds = gpu.DeviceArray(128 * 4096 * 4)
    for s0 in range(0, n, 128):
        sl = range(s0, min(n, s0 + 128))
        ms, gs, oys = [], [], []
        for s in sl:
            ox, oy, oz = (int(v) for v in meta["origin"][s][:3])
            m, g = gpu.fill_chunk(p, ox // 64, oy // 64, oz // 64)
            ms.append(m)
            gs.append(g)
            oys.append(oy)
        mats.upload(np.concatenate(ms))
        grounds.upload(np.concatenate(gs))
        counts = b.count(mats, grounds, oys, p.dirt_depth)
        assert np.array_equal(counts, meta["node_count"][s0:s0 + len(sl)])
        b.emit(meta["node_base"][s0:s0 + len(sl)].astype(np.uint64), meta["macro_base"][s0:s0 + len(sl)],
               meta["colinfo_base"][s0:s0 + len(sl)])
    assert np.array_equal(ctx.read_region(gpu.ARENA_NODES, 0, len(hw.nodes)).view(np.uint32),
                          hw.nodes.view(np.uint32))
    assert np.array_equal(ctx.read_region(gpu.ARENA_ROPES, 0, len(hw.ropes)).view(np.uint32),
                          hw.ropes.view(np.uint32))
    assert np.array_equal(ctx.read_region(gpu.ARENA_MACRO, 0, n * 16), hw.macro)
    assert np.array_equal(ctx.read_region(gpu.ARENA_COLINFO, 0, n * 2048), hw.colinfo)
    b.close()
    ctx.close()
