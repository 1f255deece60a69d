"""The C++ chunk builder (product) against the numpy restatement of the
reference builder (oracle/svo_builder.py), plus the structural invariants the
reference asserts (svo/builder_svo.rs:24-26,435,497; builder_prefix.rs:47)
and geometric checks of ropes, macro bits and column tops."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import host_world

INVALID = 0xFFFFFFFF


def _chunks_of_interest(vx):
    p = vx.world_params(42)
    # surface chunks (terrain + trees), a fully underground chunk, an empty sky chunk
    return [(p, 3, 1, 2), (p, 0, 1, 0), (p, 5, 2, 6), (p, 1, 0, 1), (p, 2, 3, 2)]


@pytest.mark.parametrize("idx", range(5))
def test_builder_matches_restatement(vx, idx):
    from oracle import svo_builder as ref

    p, cx, cy, cz = _chunks_of_interest(vx)[idx]
    mat, ground = vx.fill_chunk(p, cx, cy, cz)
    got = vx.build_chunk(mat, ground, cy * 64, p.dirt_depth)
    n, r, mw, ci = ref.build_chunk(mat, ground, cy * 64, p.dirt_depth)
    assert np.array_equal(got.nodes.view(np.uint32).reshape(-1, 4), n)
    assert np.array_equal(got.ropes.view(np.uint32).reshape(-1, 8), r)
    assert np.array_equal(got.macro, mw)
    assert np.array_equal(got.colinfo, ci)


def _random_volume(rng, density, with_ground=True):
    mat = np.zeros((64, 64, 64), np.uint8)
    # a few random boxes of random materials plus sparse noise
    for _ in range(rng.integers(1, 12)):
        x0, y0, z0 = rng.integers(0, 60, 3)
        sx, sy, sz = rng.integers(1, 24, 3)
        mat[y0:y0 + sy, z0:z0 + sz, x0:x0 + sx] = rng.integers(1, 6)
    noise = rng.random((64, 64, 64)) < density
    mat[noise] = rng.integers(1, 7, noise.sum())
    ground = rng.integers(0, 128, (64, 64)).astype(np.int32) if with_ground else np.full((64, 64), -1000, np.int32)
    return mat.reshape(-1), ground.reshape(-1)


@pytest.mark.parametrize("seed", range(4))
def test_builder_random_volumes(vx, seed):
    from oracle import svo_builder as ref

    rng = np.random.default_rng(seed)
    mat, ground = _random_volume(rng, density=[0.0, 0.001, 0.02, 0.2][seed])
    got = vx.build_chunk(mat, ground, 0, 4)
    n, r, mw, ci = ref.build_chunk(mat, ground, 0, 4)
    assert np.array_equal(got.nodes.view(np.uint32).reshape(-1, 4), n)
    assert np.array_equal(got.ropes.view(np.uint32).reshape(-1, 8), r)
    assert np.array_equal(got.macro, mw)
    assert np.array_equal(got.colinfo, ci)


def _cube(key):
    lvl = (key >> 18) & 7
    s = 64 >> lvl
    return np.array([(key & 63) * s, ((key >> 6) & 63) * s, ((key >> 12) & 63) * s]), s


def _walk(nodes):
    """Reachable nodes from the root: (index, parent, ci)."""
    out, stack = [], [0]
    while stack:
        i = stack.pop()
        out.append(i)
        cb, m = int(nodes[i]["child_base"]), int(nodes[i]["child_mask"])
        if cb == INVALID:
            continue
        rank = 0
        for ci in range(8):
            if m >> ci & 1:
                stack.append(cb + rank)
                rank += 1
    return out


def test_invariants_on_seed42_chunks(vx):
    sc, hw = host_world("c0")
    p = sc.params
    for slot in (37, 74, 111, 120, 99):
        m = hw.meta[slot]
        base, cnt = int(m["node_base"]), int(m["node_count"])
        nodes = hw.nodes[base:base + cnt]
        ropes = hw.ropes[base:base + cnt]
        cx, cy, cz = (int(v) // 64 for v in m["origin"][:3])
        mat, ground = vx.fill_chunk(p, cx, cy, cz)
        mat3 = mat.reshape(64, 64, 64)
        # root first, levels contiguous and child ranges in bounds
        assert (int(nodes[0]["key"]) >> 18) & 7 == 0
        levels = (nodes["key"] >> 18) & 7
        assert np.all(np.diff(levels.astype(np.int64)) >= 0)
        reach = _walk(nodes)
        for i in reach:
            cb = int(nodes[i]["child_base"])
            if cb != INVALID:
                k = bin(int(nodes[i]["child_mask"])).count("1")
                assert 0 < k and cb + k <= cnt
        # leaves tile the solid voxels exactly
        occ = np.zeros_like(mat3, bool)
        for i in reach:
            if int(nodes[i]["child_base"]) != INVALID:
                continue
            if int(nodes[i]["material"]) == 0:
                continue
            (x, y, z), s = _cube(int(nodes[i]["key"]))
            assert not occ[y:y + s, z:z + s, x:x + s].any()
            occ[y:y + s, z:z + s, x:x + s] = True
            sub = mat3[y:y + s, z:z + s, x:x + s]
            if int(nodes[i]["material"]) != 3 or s == 1:  # stone collapse may merge non-stone
                assert np.all(sub == nodes[i]["material"])
            assert np.all(sub != 0)
        assert np.array_equal(occ, mat3 != 0)
        # macro bits = OR over each 8^3 cell; macro_empty <=> all zero
        mw = hw.macro[slot * 16:(slot + 1) * 16]
        cells = (mat3 != 0).reshape(8, 8, 8, 8, 8, 8).any(axis=(1, 3, 5))  # [my, mz, mx]
        for mz in range(8):
            for my in range(8):
                for mx in range(8):
                    bit = mx + 8 * (my + 8 * mz)
                    assert bool(int(mw[bit >> 5]) >> (bit & 31) & 1) == bool(cells[my, mz, mx])
        assert bool(m["macro_empty"]) == (not mw.any())
        # column tops = max solid y, material of that voxel
        ci = hw.colinfo[slot * 2048:(slot + 1) * 2048]
        e = np.empty(4096, np.uint32)
        e[0::2], e[1::2] = ci & 0xFFFF, ci >> 16
        e = e.reshape(64, 64)  # [z, x]
        solid = mat3 != 0
        for z in range(0, 64, 7):
            for x in range(0, 64, 5):
                col = np.nonzero(solid[:, z, x])[0]
                if len(col) == 0:
                    assert e[z, x] & 0xFF == 255
                else:
                    assert e[z, x] & 0xFF == col[-1]
                    assert e[z, x] >> 8 == mat3[col[-1], z, x]
        # ropes (builder_ropes.rs:106-173): a target lies beyond the face, is of
        # the same or a coarser level and covers the node's face footprint.  It
        # is face adjacent when the sibling across the face exists; when that
        # sibling is missing the reference uses the parent's rope descended one
        # level, so the target can lie further along the axis.  Whatever it
        # skips (between the face and the target, over the footprint) is air.
        for i in reach[1:]:
            (o, s) = _cube(int(nodes[i]["key"]))
            r = ropes[i]
            for f, name in enumerate(("px", "nx", "py", "ny", "pz", "nz")):
                t = int(r[name])
                if t == INVALID:
                    continue
                (ot, st) = _cube(int(nodes[t]["key"]))
                assert st >= s
                ax, pos = f // 2, f % 2 == 0
                lo, hi = (o[ax] + s, ot[ax]) if pos else (ot[ax] + st, o[ax])
                assert lo <= hi and (hi - lo) % s == 0
                sl = [slice(o[b], o[b] + s) for b in range(3)]
                sl[ax] = slice(lo, hi)
                assert not mat3[sl[1], sl[2], sl[0]].any()  # skipped cells are air
                for b in range(3):
                    if b != ax:
                        assert ot[b] <= o[b] and o[b] + s <= ot[b] + st


def test_world_grid_and_meta(vx):
    sc, hw = host_world("c0")
    nx, ny, nz = hw.grid_dims
    assert len(hw.grid) == nx * ny * nz and hw.chunk_count == nx * ny * nz
    for slot in range(hw.chunk_count):
        m = hw.meta[slot]
        assert m["macro_base"] == slot * 16 and m["colinfo_base"] == slot * 2048
        ix, iy, iz = (int(v) // 64 for v in m["origin"][:3])
        assert hw.grid[(iz * ny + iy) * nx + ix] == slot
    # node ranges are contiguous and disjoint
    ends = hw.meta["node_base"].astype(np.int64) + hw.meta["node_count"]
    assert np.array_equal(hw.meta["node_base"][1:].astype(np.int64), ends[:-1]) and ends[-1] == len(hw.nodes)


def test_world_generator_spec(vx):
    """Seed-42 world: ground in [48, 144], stone below ground-4, dirt, grass top, ~1% tree columns."""
    p = vx.world_params(42)
    g = np.array([[vx.ground_height(p, x, z) for x in range(0, 512, 3)] for z in range(0, 512, 3)])
    assert g.min() >= 48 and g.max() <= 144 and g.std() > 5
    mat, ground = vx.fill_chunk(p, 2, 1, 5)
    mat3 = mat.reshape(64, 64, 64)
    trees = 0
    for z in range(64):
        for x in range(64):
            gh = int(ground[z * 64 + x]) - 64
            col = mat3[:, z, x]
            for y in range(64):
                if y < gh - 4:
                    assert col[y] == 3
                elif y < gh:
                    assert col[y] == 2
                elif y == gh:
                    assert col[y] == 1
            if 0 <= gh + 1 < 64 and col[gh + 1] == 4:
                trees += 1
    assert 10 <= trees <= 90  # ~1% of 4096 columns carry a trunk start in this chunk


def test_edit_rebuilds_one_chunk(vx):
    p = vx.world_params(42)
    hw = vx.HostWorld(p, (0, 0, 0), (2, 2, 2))
    before_nodes = hw.nodes.copy()
    slot_meta = hw.meta.copy()
    # dig a hole at a surface voxel of chunk (1, 1, 0)
    x, z = 64 + 10, 20
    y = vx.ground_height(p, x, z)
    assert 64 <= y < 128
    rebuilt = hw.edit(np.array([[x, y, z]]), np.array([0]))
    assert rebuilt == [1 + 2 * 1]  # slot of chunk (1, 1, 0): grid order, x fastest
    mat, ground = vx.fill_chunk(p, 1, 1, 0)
    mat[(y - 64) * 4096 + z * 64 + (x - 64)] = 0
    ref = vx.build_chunk(mat, ground, 64, p.dirt_depth)
    m = hw.meta[3]
    got = hw.nodes[int(m["node_base"]):int(m["node_base"]) + int(m["node_count"])]
    assert np.array_equal(got.view(np.uint32), ref.nodes.view(np.uint32))
    # other chunks unchanged
    for other in (0, 1, 2):
        a = slot_meta[other]
        b = hw.meta[other]
        assert np.array_equal(before_nodes[int(a["node_base"]):int(a["node_base"]) + int(a["node_count"])].view(np.uint32),
                              hw.nodes[int(b["node_base"]):int(b["node_base"]) + int(b["node_count"])].view(np.uint32))
