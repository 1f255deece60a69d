"""Pins the oracle's sky, godray march and surface shading to float64
statements of the reference WGSL (tests/spec64.py), on worlds built so that
every input of the statement is known independently of the oracle:

  * the cloud march of the sky (sky.wgsl:43-137, clouds.wgsl:35-121, fbm and
    value noise of common.wgsl:614-662) on an empty world, cameras under,
    inside and above the cloud slab;
  * the godray march with fog on (godrays.wgsl:22-216, 241-356): on an empty
    world (no geometry: the shaft terms are zero) and under a roof of
    full-chunk stone slabs with holes, where the geometry-only sun
    transmittance is 0 or 1 and the shaft low-pass, edge boost and sun gate
    are live;
  * the shading of dirt, stone and wood voxel faces (shading.wgsl:7-52,
    137-362, ray_main.wgsl:428-470): material variation, macro AO taps, the
    blended sun shadow, cloud shadow, ambient, Blinn-Phong specular with
    Schlick Fresnel, fog, and the unfogged local-light output;
  * the shading of grass heightfield hits on a flat clipmap below the wind
    ramp (clipmap.wgsl:156-275, shading.wgsl:52-135, 364-434): the trace with
    its three bisection steps, clip material variation, the grass-top albedo
    (blade slivers, fbm patches) and the terrain lighting.

A misread line (a transposed rotation, a wrong constant, a dropped term, a
flipped comparison) changes the compared values by far more than the
tolerances used here; the module docstring of each test names what is live.

Two behaviours of the reference are modelled here explicitly, because the
geometry-only sun walk depends on them (shadows.wgsl:5-80, svo_query.wgsl:11-55):
  * a leaf query in an empty 8^3 macro cell, or outside the chunk cube,
    returns the whole chunk as one AIR leaf, so the walk leaves the chunk at
    once: a sun ray that enters a chunk through an empty macro cell sees
    nothing else in that chunk;
  * the walk starts 0.02 m (0.2 voxel) past the chunk's entry and queries
    nothing when that start already lies past the chunk's exit.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import spec64 as S

VS = 0.1          # voxel size (m)
CS = 6.4          # chunk size (m)


@pytest.fixture(scope="module")
def O():
    from oracle import oracle

    return oracle


@pytest.fixture(scope="module")
def vx():
    import voxhip

    return voxhip


def _empty_world(O, vx, clip=None, heights=None):
    if clip is None:
        clip, heights = vx.Clipmap(), np.zeros((1, 1, 1), np.float32)
    return O.World(np.zeros(1, vx.NODE_DTYPE), np.zeros(1, vx.ROPES_DTYPE), np.zeros(1, vx.META_DTYPE),
                   np.full(1, 0xFFFFFFFF, np.uint32), np.zeros(16, np.uint32), np.zeros(2048, np.uint32), clip, heights)


def _pixel_rays(cam, rw, rh):
    y, x = np.mgrid[0:rh, 0:rw]
    x, y = x.ravel(), y.ravel()
    return x, y, S.ray_dir(cam, x + 0.5, y + 0.5), np.array(cam.cam_pos[:3], np.float64)


SUN_YAW = math.atan2(S.SUN_DIR[0], S.SUN_DIR[2])


# ---------------------------------------------------------------- sky with clouds
@pytest.mark.parametrize("pos,dyaw,pitch,time_s", [
    ((5.0, 20.0, 7.0), 0.0, 0.4, 0.0),            # under the slab, towards the sun
    ((300.0, 60.0, -200.0), 1.3, 0.8, 37.5),      # under the slab, steep
    ((-50.0, 169.0, 40.0), 2.5, 0.2, 500.0),      # just under the base
    ((0.0, 200.0, 0.0), 0.5, 0.1, 12.0),          # inside the slab
])
def test_sky_cloud_march(O, vx, pos, dyaw, pitch, time_s):
    """Empty world: every primary pixel is sky_color(rd), here with the
    8-step cloud march, the 4-step self-shadow, the silver lining and the
    sun-disc dimming live.  Rays where a float32 march could take the other
    side of `dens > 1e-4` or `T < 0.02`, and the sun disc's edge, are left
    out."""
    rw, rh = 128, 80
    w = _empty_world(O, vx)
    cam = vx.camera_build(vx.make_pose(pos, SUN_YAW + dyaw, pitch, 70.0), (rw, rh), (rw, rh), 0, chunk_count=0,
                          time_s=time_s)
    fr = O.Frame(rw, rh)
    O.primary(w, cam, fr)
    _, _, rd, ro = _pixel_rays(cam, rw, rh)
    m = {}
    ref = S.sky_color(rd, ro, float(cam.voxel_params[1]), m)
    mu = rd @ S.SUN_DIR
    edge = (mu > S.SUN_MU_OUTER - 2e-5) & (mu < S.SUN_MU_INNER + 2e-5)
    cmp = ~edge & ~m["near"]
    got = fr.color.reshape(-1, 4)[:, :3]
    assert cmp.mean() > 0.98
    err = np.abs(got - ref).max(-1)[cmp]
    assert err.max() <= 1e-4, err.max()
    cloudy = m["T"][cmp] < 0.999
    assert cloudy.mean() > 0.03            # the march adds radiance on a share of the compared rays


# ---------------------------------------------------------------- godray march, fog on
def _godray_compare(O, vx, w, cam, depth, ts_geom=None, macro_compat=True, amb=None):
    rh, rw = depth.shape
    fr = O.Frame(rw, rh)
    fr.depth[...] = depth
    O.godray(w, cam, fr, macro_compat=macro_compat)
    inp = S.godray_pixel_inputs(cam, depth)
    ro = np.array(cam.cam_pos[:3], np.float64)
    g, near = S.godray_integrate(ro, inp["rd"], inp["t_end"], inp["j4"], inp["j_phase"], inp["steps"],
                                 float(cam.voxel_params[3]), float(cam.voxel_params[1]), ts_geom=ts_geom)
    ref = S.godray_compress(np.maximum(g, 0.0))
    cmp = ~near & ~inp["near"]
    if amb is not None:
        cmp &= ~amb
    got = fr.godray.reshape(-1, 4)
    # alpha: the history depth proxy, the minimum of the five depth taps
    d = np.pad(depth, 1, mode="edge")
    dmin = np.minimum.reduce([d[1:-1, 1:-1], d[:-2, 1:-1], d[2:, 1:-1], d[1:-1, :-2], d[1:-1, 2:]]).ravel()
    assert np.all(got[:, 3] == dmin)
    return np.abs(got[:, :3] - ref).max(-1), cmp, ref, np.abs(inp["rd"][:, 1])


def _assert_godray_close(err, cmp, dy):
    """1e-5 where the view ray climbs or falls by at least 1 %; near-horizontal
    rays get 1e-3: there the fog recurrence's base * (B - B_s) / (k dy)
    cancels in float32, as the reference writes it."""
    steep = dy >= 0.01
    assert err[cmp & steep].max() <= 1e-5, err[cmp & steep].max()
    if (cmp & ~steep).any():
        assert err[cmp & ~steep].max() <= 1e-3, err[cmp & ~steep].max()


@pytest.mark.parametrize("pos,dyaw,pitch,time_s,fog,frame", [
    ((5.0, 2.0, 7.0), 0.3, 0.1, 3.0, 0.02, 5),
    ((300.0, 30.0, -200.0), 2.0, -0.3, 37.5, 0.05, 77),
    ((0.0, 10.0, 0.0), 0.0, 0.3, 120.0, 0.01, 1000),
])
def test_godray_march_empty_world(O, vx, pos, dyaw, pitch, time_s, fog, frame):
    """No geometry: the geometry-only sun transmittance is 1, so the shaft
    terms vanish and the march is the height-fog transmittance recurrence,
    the haze ramp, the scattering height term, phase, view gate, the cloud
    term of the sun ray, the step count and jitters.  No history is valid
    (the history depth is zero), so the output is compress(current)."""
    rw, rh = 96, 64
    w = _empty_world(O, vx)
    cam = vx.camera_build(vx.make_pose(pos, SUN_YAW + dyaw, pitch, 70.0), (rw, rh), (rw, rh), frame, chunk_count=0,
                          time_s=time_s, fog_density=fog)
    depth = np.random.default_rng(3).uniform(0.5, 130.0, (rh, rw)).astype(np.float32)
    err, cmp, ref, dy = _godray_compare(O, vx, w, cam, depth)
    assert cmp.mean() > 0.99
    _assert_godray_close(err, cmp, dy)
    assert ref.mean() > 0.01


def _roof_world(vx, O, dims=(8, 2, 8), full=False):
    """Full-chunk stone slabs at voxel rows 32..40 (9.6..10.4 m) of the upper
    chunk layer, in a pattern of chunks with holes (or in every chunk)."""
    import worlds

    roof = set()
    vols = {}
    for cx in range(dims[0]):
        for cz in range(dims[2]):
            if full or (cx + cz) % 2 == 0 or (cx * 7 + cz * 3) % 5 == 0:
                m = np.zeros((64, 64, 64), np.uint8)
                m[32:40, :, :] = 3
                vols[(cx, 1, cz)] = m.reshape(-1)
                roof.add((cx, 1, cz))
    pw = worlds.pack(vx, dims, vols)
    clip, heights = worlds.empty_clipmap(vx)
    return pw, O.World(pw.nodes, pw.ropes, pw.meta, pw.grid, pw.macro, pw.colinfo, clip, heights), roof


def roof_sun_transmittance(p, roof, dims, slab_y=(9.6, 10.4), margin=2e-3):
    """sun_transmittance_geom_only (shadows.wgsl:83-185) over a roof world,
    in float64: ro = p + 0.05 sun; the walk starts 0.018 m inside the grid;
    each chunk on the chunk DDA is queried 0.02 m past its entry (nothing is
    queried when that lies past its exit); the query is solid exactly when it
    lies in a roof chunk's slab rows, and any other query returns the whole
    chunk as air.  Returns (transmittance, ambiguous) per point."""
    d = S.SUN_DIR
    gmax = np.array(dims, float) * CS
    out = np.ones(len(p))
    amb = np.zeros(len(p), bool)
    for k in range(len(p)):
        ro = p[k] + d * 0.05
        t0, t1 = (0.0 - ro) / d, (gmax - ro) / d
        te, tx = max(np.minimum(t0, t1).max(), 0.0), np.maximum(t0, t1).min()
        if tx < te:
            continue
        start = te + 0.018
        cr = []
        for a in range(3):
            c0 = np.floor((ro[a] + start * d[a]) / CS)
            for m in range(1, 64):
                t = ((c0 + m) * CS - ro[a]) / d[a]
                if t > tx:
                    break
                cr.append(t)
        seq = [start] + sorted(cr)
        if len(seq) > 1 and np.diff(seq).min() < margin:
            amb[k] = True
        for j, tin in enumerate(seq):
            tout = min(seq[j + 1] if j + 1 < len(seq) else tx, tx)
            tq = tin + 0.02
            amb[k] |= abs(tq - tout) < margin
            if tq > tout:
                continue
            c = tuple(np.floor((ro + (tin + 1e-9) * d) / CS).astype(int))
            if any(ci < 0 or ci >= di for ci, di in zip(c, dims)):
                break
            if c in roof:
                q = ro + tq * d
                amb[k] |= min(abs(q[1] - slab_y[0]), abs(q[1] - slab_y[1])) < margin
                if slab_y[0] <= q[1] < slab_y[1]:
                    out[k] = 0.0
                    break
    return out, amb


@pytest.mark.parametrize("full", [False, True])
def test_godray_march_shafts_under_a_roof(O, vx, full):
    """Cameras under a roof of stone slabs: the sun transmittance along the
    march switches between 0 and 1, so ts, the shaft low-pass, its edge
    boost and the sun gate all move.  The pass is run with the macro words
    bound as the chunk's own (the H2 fix); the compat binding differs only in
    which words the leaf query's macro test reads."""
    pw, w, roof = _roof_world(vx, O, full=full)
    rw, rh = 112, 64
    rng = np.random.default_rng(5)
    n_shadow = [0, 0]
    for pos, yaw, pitch, frame, time_s, fog in [((20.0, 2.0, 22.0), 0.6, 0.15, 0, 10.0, 0.05),
                                                 ((30.0, 4.0, 12.0), -0.3, -0.1, 3, 200.0, 0.02)]:
        cam = vx.camera_build(vx.make_pose(pos, yaw, pitch, 70.0), (rw, rh), (rw, rh), frame, time_s=time_s,
                              fog_density=fog, grid_origin=(0, 0, 0), grid_dims=(8, 2, 8), chunk_count=pw.chunk_count)
        depth = rng.uniform(5.0, 110.0, (rh, rw)).astype(np.float32)
        amb = np.zeros(rw * rh, bool)

        def ts_geom(p, idx):
            tr, a = roof_sun_transmittance(p, roof, (8, 2, 8))
            amb[idx] |= a
            n_shadow[0] += int((tr == 0.0).sum())
            n_shadow[1] += len(tr)
            return tr

        err, cmp, _, dy = _godray_compare(O, vx, w, cam, depth, ts_geom=ts_geom, macro_compat=False, amb=amb)
        assert cmp.mean() > 0.95
        _assert_godray_close(err, cmp, dy)
    assert 0.05 < n_shadow[0] / n_shadow[1] < 0.95      # shadowed and lit samples both occur


# ---------------------------------------------------------------- voxel surface shading
SLABS = [  # chunk (cx, cz) in the lower layer, material, voxel box x0, x1, y0, y1, z0, z1 inside it
    ((1, 1), 3, (2, 24, 10, 14, 3, 60)),
    ((2, 1), 2, (5, 60, 20, 26, 2, 50)),
    ((1, 2), 4, (3, 56, 12, 20, 4, 61)),
    ((2, 2), 3, (4, 60, 30, 33, 4, 60)),
    ((3, 3), 4, (6, 58, 16, 22, 6, 58)),
    ((4, 4), 2, (2, 62, 8, 12, 2, 62)),
    ((6, 6), 3, (2, 62, 8, 40, 2, 62)),
    ((8, 8), 2, (2, 62, 8, 40, 2, 62)),
    ((10, 10), 4, (2, 62, 8, 40, 2, 62)),
    ((3, 1), 3, (2, 62, 8, 24, 2, 62)),
    ((1, 3), 2, (2, 62, 8, 24, 2, 62)),
    ((5, 3), 4, (2, 62, 8, 24, 2, 62)),
    ((3, 5), 3, (2, 62, 8, 24, 2, 62)),
]
SLAB_DIMS = (12, 2, 12)


@pytest.fixture(scope="module")
def slab_world(vx, O):
    import worlds

    vols, boxes = {}, []
    for (cx, cz), mat, (x0, x1, y0, y1, z0, z1) in SLABS:
        m = np.zeros((64, 64, 64), np.uint8)
        m[y0:y1, z0:z1, x0:x1] = mat
        vols[(cx, 0, cz)] = m.reshape(-1)
        lo = np.array([cx * 64 + x0, y0, cz * 64 + z0], np.float64) * VS
        hi = np.array([cx * 64 + x1, y1, cz * 64 + z1], np.float64) * VS
        boxes.append((lo, hi, mat, (cx, cz), m))
    pw = worlds.pack(vx, SLAB_DIMS, vols)
    clip, heights = worlds.empty_clipmap(vx)
    return pw, O.World(pw.nodes, pw.ropes, pw.meta, pw.grid, pw.macro, pw.colinfo, clip, heights), boxes


def _ray_boxes(ro, rd, boxes):
    n = len(rd)
    best, bi, ax = np.full(n, np.inf), np.full(n, -1), np.full(n, -1)
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = 1.0 / rd
    for k, (lo, hi, *_) in enumerate(boxes):
        t0, t1 = (lo[None, :] - ro[None, :]) * inv, (hi[None, :] - ro[None, :]) * inv
        tmin, tmax = np.minimum(t0, t1), np.maximum(t0, t1)
        te, tx = tmin.max(-1), tmax.min(-1)
        hit = (te <= tx) & (te > 0) & (te < best)
        best[hit], bi[hit], ax[hit] = te[hit], k, tmin[hit].argmax(-1)
    return best, bi, ax


def _macro_occ(boxes, chunk, q):
    """Macro occupancy (8^3 cells) of the hit chunk at points q, 0 outside its
    cube, with a flag for points within 8e-5 m of a cell or chunk face."""
    vol = {c: mm for (_, _, _, c, mm) in boxes}
    lp = q - np.stack([chunk[:, 0] * CS, np.zeros(len(q)), chunk[:, 1] * CS], -1)
    inside = np.all((lp >= 0) & (lp < CS), -1)
    f = lp / 0.8 - np.floor(lp / 0.8)
    amb = np.any(np.minimum(f, 1.0 - f) < 1e-4, -1)
    cell = np.floor(lp / 0.8).astype(int).clip(0, 7)
    occ = np.zeros(len(q))
    for i in np.nonzero(inside)[0]:
        mm = vol[(int(chunk[i, 0]), int(chunk[i, 1]))]
        mx, my, mz = cell[i]
        occ[i] = float(mm[my * 8:(my + 1) * 8, mz * 8:(mz + 1) * 8, mx * 8:(mx + 1) * 8].any())
    return occ, amb


def _sun_blocked(boxes, p, margin):
    d = S.SUN_DIR
    out = np.zeros(len(p), bool)
    for lo, hi, *_ in boxes:
        t0, t1 = (lo - margin - p) / d, (hi + margin - p) / d
        te, tx = np.minimum(t0, t1).max(-1), np.maximum(t0, t1).min(-1)
        out |= (te <= tx) & (tx > 0)
    return out


@pytest.mark.parametrize("pos,yaw,pitch,frame,time_s,fog", [
    ((2.0, 9.0, 2.0), math.pi / 4, -0.45, 0, 10.0, 0.05),
    ((30.0, 12.0, 2.0), -0.3, -0.6, 3, 200.0, 0.02),
])
def test_voxel_face_shading(O, vx, slab_world, pos, yaw, pitch, frame, time_s, fog):
    """Faces of dirt, stone and wood slabs floating in an empty world, shadow
    history 1: shade_hit_split + apply_fog in float64 from the float64
    ray/box hit (the oracle's depth must agree to 1e-5).  Left out: hits whose
    face is decided by the traversal's 1e-3-voxel nudges (within 3e-4 m of a
    voxel edge), hits where an AO tap, a material-variation cell or a shadow
    ray is within float32 reach of a boundary.  The variation cell along the
    face normal sits exactly on a cell boundary (faces lie on multiples of
    0.1 m, cells are 0.05 m), where float32 rounding of the hit point picks
    the cell: either of the two is accepted there."""
    pw, w, boxes = slab_world
    rw, rh = 192, 108
    cam = vx.camera_build(vx.make_pose(pos, yaw, pitch, 70.0), (rw, rh), (rw, rh), frame, time_s=time_s,
                          fog_density=fog, grid_origin=(0, 0, 0), grid_dims=SLAB_DIMS, chunk_count=pw.chunk_count)
    fr = O.Frame(rw, rh)
    fr.shadow[...] = 1.0
    O.primary(w, cam, fr, vx.HIT_DTYPE)
    x, y, rd, ro = _pixel_rays(cam, rw, rh)
    t, bi, ax = _ray_boxes(ro, rd, boxes)
    hits, dep = fr.hits.ravel(), fr.depth.ravel().astype(np.float64)
    sel = np.isfinite(t) & (hits["kind"] == 1)
    sel[sel] &= np.abs(dep[sel] - t[sel]) <= 1e-5 * t[sel]
    assert sel.sum() > 1500
    idx = np.nonzero(sel)[0]
    k_ = np.arange(len(idx))
    rd, t, a = rd[idx], t[idx], ax[idx]
    hp = ro + rd * t[:, None]
    hp_o = ro + rd * dep[idx][:, None]
    n = np.zeros((len(idx), 3))
    n[k_, a] = -np.sign(rd[k_, a])
    mats = np.array([boxes[b][2] for b in bi[idx]])
    assert np.all(hits["mat"][idx] == mats)
    chunk = np.array([boxes[b][3] for b in bi[idx]])
    inplane = np.arange(3)[None, :] != a[:, None]
    qv = hp / VS
    amb = np.any(inplane & (np.abs(qv - np.rint(qv)) * VS < 3e-4), -1)
    # macro AO: four taps 0.75 voxel along the face's tangent frame
    up_ref = np.where((np.abs(n[:, 1]) > 0.9)[:, None], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0])
    tt = np.cross(up_ref, n)
    tt /= np.linalg.norm(tt, axis=-1, keepdims=True)
    bb = np.cross(n, tt)
    bb /= np.linalg.norm(bb, axis=-1, keepdims=True)
    occ = np.zeros(len(idx))
    for st, sb in ((1, 0), (-1, 0), (0, 1), (0, -1)):
        o1, a1 = _macro_occ(boxes, chunk, hp + (st * tt + sb * bb) * 0.75 * VS)
        o2, a2 = _macro_occ(boxes, chunk, hp_o + (st * tt + sb * bb) * 0.75 * VS)
        occ += o1
        amb |= a1 | a2 | (o1 != o2)
    ao = np.clip(1.0 - 0.70 * occ * 0.25, 0.35, 1.0)
    # one pixel in eight blends the geometric sun term into the history (0.2)
    seed = (x[idx] * 1973) ^ (y[idx] * 9277) ^ (frame * 26699)
    do = (seed & 7) == 0
    ps = hp + n * 0.75 * VS + S.SUN_DIR * 0.068
    grow, shrink = _sun_blocked(boxes, ps, 0.03), _sun_blocked(boxes, ps, -0.03)
    amb |= do & (grow != shrink)
    cur = np.where(shrink, 0.0, 1.0)
    shadow = np.where(do, 0.8 + 0.2 * cur, 1.0)
    # material variation cells
    q, q_o = hp / 0.05, hp_o / 0.05
    cells = np.floor(q).astype(np.int64)
    f = q - np.floor(q)
    amb |= np.any(inplane & ((np.minimum(f, 1.0 - f) < 5e-4) | (np.floor(q_o) != np.floor(q))), -1)
    face = np.rint(q[k_, a]).astype(np.int64)
    got = fr.color.reshape(-1, 4)[idx, :3].astype(np.float64)
    sky = S.sky_bg(rd)
    err = np.full(len(idx), np.inf)
    for choice in (face, face - 1):
        cl = cells.copy()
        cl[k_, a] = choice
        col = np.zeros((len(idx), 3))
        for m in (2, 3, 4):
            s = mats == m
            alb = S.material_albedo(m, cl[s], t[s])
            hdr = S.shade_voxel(ro, rd[s], t[s], n[s], m, alb, ao[s], shadow[s], time_s)
            col[s] = S.apply_fog(hdr, ro, rd[s], np.minimum(t[s], S.FOG_MAX_DIST), sky[s], fog)
        err = np.minimum(err, np.abs(col - got).max(-1))
    ok = ~amb
    assert ok.mean() > 0.5
    assert err[ok].max() <= 1e-5, err[ok].max()
    # branches live among the compared hits
    assert set(np.unique(mats[ok])) == {2, 3, 4}
    assert ((ao[ok] > 0.35) & (ao[ok] < 1.0)).any() and (do & ok & (cur == 0.0)).any() and (do & ok & (cur == 1.0)).any()
    assert (np.abs(n[ok, 1]) < 0.5).any() and (n[ok, 1] > 0.5).any()
    if pos[1] > 10.0:
        assert (t[ok] > 64.0).any() and (t[ok] > 40.0).any()
    # local output: no lights, so zero radiance, valid (w = 1) within 50 m
    loc = fr.local.reshape(-1, 4)[idx]
    assert np.all(loc[:, :3] == 0.0) and np.all(loc[:, 3] == (t <= 50.0))


# ---------------------------------------------------------------- heightfield grass shading
H_FLAT, CELL = 1.234, 4.0


def _flat_clipmap_world(O, vx, cx, cz):
    clip = vx.Clipmap()
    clip.levels, clip.res, clip.base_cell_m = 1, 256, CELL
    clip.level[0][0] = float(np.floor(cx / CELL) * CELL - 128 * CELL)
    clip.level[0][1] = float(np.floor(cz / CELL) * CELL - 128 * CELL)
    clip.level[0][2] = CELL
    return _empty_world(O, vx, clip, np.full((1, 256, 256), H_FLAT, np.float32))


@pytest.mark.parametrize("pos,dyaw,pitch,time_s,fog", [
    ((5.0, 4.0, 7.0), 0.5, -0.35, 3.0, 0.03),
    ((-40.0, 9.0, 25.0), 2.9, -0.6, 50.0, 0.01),
    ((100.0, 2.5, -30.0), 1.0, -0.2, 0.0, 0.05),
])
def test_heightfield_grass_shading(O, vx, pos, dyaw, pitch, time_s, fog):
    """No chunks, one clip level of constant height 1.234 m: every downward
    ray hits the heightfield.  The hit distances the trace can produce are
    restated in float64 (its vertical step lands on the plane, where float32
    rounding picks one of two outcomes); every oracle heightfield hit must be
    one of them unless a grass blade of the slab test was hit first (then the
    depth is the blade's).  Below the wind ramp the blade direction is fixed,
    so the grass-top albedo is a closed form of hashes and fbm."""
    rw, rh = 160, 96
    w = _flat_clipmap_world(O, vx, pos[0], pos[2])
    cam = vx.camera_build(vx.make_pose(pos, SUN_YAW + dyaw, pitch, 70.0), (rw, rh), (rw, rh), 0, chunk_count=0,
                          time_s=time_s, fog_density=fog)
    fr = O.Frame(rw, rh)
    O.primary(w, cam, fr, vx.HIT_DTYPE)
    _, _, rd, ro = _pixel_rays(cam, rw, rh)
    cands = S.clip_trace_flat(ro, rd, H_FLAT, CELL)
    hit = np.array([len(c) > 0 for c in cands])
    kind = fr.hits.ravel()["kind"]
    assert np.array_equal(kind == 2, hit)
    dep = fr.depth.ravel().astype(np.float64)
    th = np.full(len(rd), np.nan)
    for k, c in enumerate(cands):
        for v in c:
            if abs(dep[k] - min(v, S.FOG_MAX_DIST)) <= 2e-6 * max(v, 1.0):
                th[k] = v
    trace_hit = hit & np.isfinite(th)
    assert trace_hit.sum() > 0.6 * hit.sum()           # the rest hit a blade first
    idx = np.nonzero(trace_hit)[0]
    flags = {"near": np.zeros(len(idx), bool)}
    hp = ro + rd[idx] * th[idx, None]
    assert hp[:, 1].max() < 2.0                         # below WIND_RAMP_Y0: the wind field is zero
    surf = S.shade_clip_hit_flat(ro, rd[idx], th[idx], time_s, flags)
    col = S.apply_fog(surf, ro, rd[idx], np.minimum(th[idx], S.FOG_MAX_DIST), S.sky_bg(rd[idx]), fog)
    got = fr.color.reshape(-1, 4)[idx, :3]
    ok = ~flags["near"]
    assert ok.mean() > 0.9
    err = np.abs(col - got).max(-1)[ok]
    assert err.max() <= 5e-4, err.max()
    assert np.all(fr.local.reshape(-1, 4)[idx, 3] == 0.0)   # heightfield hits carry no local light
