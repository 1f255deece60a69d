"""Pins the oracle's screen-space passes to the reference's formulas on inputs
where each pass reduces to a closed form that tests/spec64.py states in
float64 from the WGSL text:

  * composite over constant fields (grade, Hable white scale, ^0.98, dither,
    gamma), an impulse in the local-light history (added after gamma, at the
    mapped render texel) and an impulse in the godray texture (read at the
    +0.5-offset texel: the reference adds half a pixel to a pixel centre);
  * local TAA on random and non-finite inputs (the mix, the confidence rule,
    the is_bad_vec3 guards);
  * composite TAA: pass-through at frame_index <= 1, and mix(hist, cur, 0.45)
    at the texel a static or moving camera reprojects to;
  * the godray history blend with fog off (current term zero): decompress,
    depth and validity gates, reactive mask, clamp, history weight, compress;
  * the sky of the primary pass where rays miss the cloud slab: background
    gradient plus sun disc, halo and glow.

A misread line of any of these formulas (a dropped term, a wrong constant,
a transposed tap, a flipped reprojection axis) changes the compared values by
far more than the tolerances used here."""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import pytest

import spec64 as S


@pytest.fixture(scope="module")
def O():
    from oracle import oracle

    return oracle


@pytest.fixture(scope="module")
def vx():
    import voxhip

    return voxhip


def _cam(vx, rw, rh, pw=None, ph=None, frame_index=3, pos=(10.0, 20.0, 30.0), yaw=0.3, pitch=-0.1, fog=0.002,
         prev=None, fovy=70.0):
    pw, ph = pw or rw, ph or rh
    pose = vx.make_pose(pos, yaw, pitch, fovy)
    return vx.camera_build(pose, (rw, rh), (pw, ph), frame_index, fog_density=fog, prev_view_proj=prev)


def _static_cam(vx, rw, rh, pw=None, ph=None, **kw):
    c0 = _cam(vx, rw, rh, pw, ph, **kw)
    return _cam(vx, rw, rh, pw, ph, prev=np.array(c0.view_proj, np.float32), **kw)


def _grid(w, h):
    y, x = np.mgrid[0:h, 0:w]
    return x.reshape(-1), y.reshape(-1)


def _run_composite(O, cam, fr):
    O.composite(cam, fr)
    return fr.pre_taa.reshape(-1, 4)


def _constant_frame(O, rw, rh, pw, ph, color, depth):
    fr = O.Frame(rw, rh, pw, ph)
    fr.color[...] = np.array([*color, 1.0], np.float32)
    fr.depth[...] = np.float32(depth)
    return fr


# ---------------------------------------------------------------- composite
CONSTANT_CASES = [
    ((0.20, 0.30, 0.40), 5.0),
    ((0.20, 0.30, 0.40), 55.0),    # godray fade and saturation boost partly on
    ((2.50, 2.00, 1.00), 10.0),    # bloom threshold crossed
    ((2.50, 2.00, 1.00), 90.0),    # bloom faded out by distance
    ((9.00, 7.00, 6.00), 35.0),    # highlight desaturation, tonemap clamp
    ((-1.0, 0.50, 0.20), 120.0),   # negative input clamped to zero
    ((0.00, 0.00, 0.00), 20.0),
]


@pytest.mark.parametrize("color,depth", CONSTANT_CASES)
def test_composite_constant_field(O, vx, color, depth):
    """Godray 0 and local history 0 over a constant colour and depth: every
    present pixel is the grade/tonemap/dither/gamma chain of that colour."""
    rw, rh = 40, 24
    cam = _cam(vx, rw, rh, frame_index=37)
    fr = _constant_frame(O, rw, rh, rw, rh, color, depth)
    got = _run_composite(O, cam, fr)
    x, y = _grid(rw, rh)
    px, py = S.px_render_from_present(x, y, rw, rh, rw, rh)
    n = len(x)
    base = np.tile(np.asarray(color, np.float64), (n, 1))
    zero = np.zeros((n, 3))
    ref = S.composite_pixel(base, [base] * 4, np.full(n, depth), zero, zero, px, py, 37, zero)
    assert np.all(got[:, 3] == 1.0)
    err = np.abs(got[:, :3] - ref)
    assert err.max() <= 2e-5, (err.max(), np.unravel_index(err.argmax(), err.shape))
    # the dither is live: neighbouring pixels of a constant field differ
    assert np.unique(got[:, 1]).size > n // 4 or np.all(got[:, :3] == got[0, :3]) and color == (0.0, 0.0, 0.0)


def test_composite_frame_index_changes_only_the_dither(O, vx):
    rw, rh = 32, 16
    fr = _constant_frame(O, rw, rh, rw, rh, (0.3, 0.3, 0.3), 12.0)
    a = _run_composite(O, _cam(vx, rw, rh, frame_index=1), fr).copy()
    b = _run_composite(O, _cam(vx, rw, rh, frame_index=257), fr).copy()   # frame_index & 255 == 1
    c = _run_composite(O, _cam(vx, rw, rh, frame_index=2), fr).copy()
    assert np.array_equal(a, b)
    d = np.abs(a[:, :3] - c[:, :3])
    # at most the dither's 1/1536 swing times gamma's slope (1/2.2) c^(1/2.2 - 1) at the graded value
    c_lin = float(np.min(a[:, :3])) ** 2.2
    assert 0.0 < d.max() <= (1.0 / 1536.0) * (1.0 / 2.2) * (c_lin - 1.0 / 1536.0) ** (1.0 / 2.2 - 1.0)


@pytest.mark.parametrize("render,present", [((32, 16), (32, 16)), ((16, 8), (32, 16)), ((24, 12), (40, 20))])
def test_composite_local_history_impulse(O, vx, render, present):
    """The local-light history is added after gamma at the render texel
    nearest to the present pixel's mapped position, and nowhere else."""
    (rw, rh), (pw, ph) = render, present
    cam = _cam(vx, rw, rh, pw, ph, frame_index=9)
    fr = _constant_frame(O, rw, rh, pw, ph, (0.4, 0.35, 0.3), 25.0)
    plain = _run_composite(O, cam, fr).copy()
    x0, y0, L = 5, 3, np.array([0.7, -0.2, 1.5], np.float32)
    fr.local_hist[y0, x0, :3] = L
    fr.local_hist[y0, x0, 3] = 1.0
    got = _run_composite(O, cam, fr)
    x, y = _grid(pw, ph)
    px, py = S.px_render_from_present(x, y, rw, rh, pw, ph)
    # nearest sample at uv = px_render / render dims: texel floor(uv * dims)
    tx = np.floor((px / np.float32(rw)) * np.float32(rw)).astype(int)
    ty = np.floor((py / np.float32(rh)) * np.float32(rh)).astype(int)
    hit = (tx == x0) & (ty == y0)
    assert hit.sum() == (pw // rw) * (ph // rh) if pw % rw == 0 else hit.sum() >= 1
    d = got[:, :3] - plain[:, :3]
    assert np.all(d[~hit] == 0.0)
    assert np.abs(d[hit] - L).max() <= 1e-6


@pytest.mark.parametrize("depth", [30.0, 60.0, 95.0])
def test_composite_godray_impulse_reads_the_offset_texel(O, vx, depth):
    """A godray impulse at texel (gx, gy) brightens present pixel (gx-1, gy-1)
    only: the shader samples at (pixel centre + 0.5) / dims (H11).  The four
    neighbour taps' depth-weighted blur sharpens the impulse's neighbours
    below zero, where the max() clamps them back to the plain value.  Past
    40 m the far fade, the sun-hue lock and the warm tint apply."""
    rw, rh = 32, 16   # power-of-two sizes: (x + 1) / 32 * 32 is exact, no sampler rounding question
    cam = _cam(vx, rw, rh, frame_index=4)
    fr = _constant_frame(O, rw, rh, rw, rh, (0.25, 0.3, 0.35), depth)
    plain = _run_composite(O, cam, fr).copy()
    gx, gy, g = 12, 7, 0.5
    fr.godray[gy, gx] = np.array([g, g, g, depth], np.float32)
    got = _run_composite(O, cam, fr)
    changed = np.any(got != plain, axis=-1).reshape(rh, rw)
    assert changed.sum() == 1 and changed[gy - 1, gx - 1], np.argwhere(changed)
    x, y = np.array([gx - 1]), np.array([gy - 1])
    px, py = S.px_render_from_present(x, y, rw, rh, rw, rh)
    base = np.array([[0.25, 0.3, 0.35]])
    g_c = np.full((1, 3), S.godray_decompress(np.float64(np.float32(g))))
    ref = S.composite_pixel(base, [base] * 4, np.array([depth]), g_c, np.zeros((1, 3)), px, py, 4, np.zeros((1, 3)))
    assert np.abs(got.reshape(rh, rw, 4)[gy - 1, gx - 1, :3] - ref[0]).max() <= 2e-5


def test_composite_godray_depth_weights(O, vx):
    """A neighbour tap across a depth edge gets weight 1 - smoothstep(0, tol,
    |dz|) = 0: with the impulse in the left neighbour tap of a pixel whose left
    depth differs by more than tol, the blur ignores it and the pixel keeps
    its unsharpened value; with equal depths the blur subtracts it."""
    rw, rh = 32, 16
    cam = _cam(vx, rw, rh, frame_index=4)
    g = 0.6
    outs = []
    for step in (0.0, 5.0):
        fr = _constant_frame(O, rw, rh, rw, rh, (0.2, 0.2, 0.2), 20.0)
        fr.depth[:, :10] = np.float32(20.0 + step)
        # pixel (10, 5): centre tap texel (11, 6), W tap texel (10, 6); depths d0 = depth[5, 10], dW = depth[5, 9]
        fr.godray[6, 11] = np.array([g, g, g, 20.0], np.float32)
        fr.godray[6, 10] = np.array([g, g, g, 20.0], np.float32)
        outs.append(_run_composite(O, cam, fr).reshape(rh, rw, 4)[5, 10, :3].copy())
    x, y = np.array([10]), np.array([5])
    px, py = S.px_render_from_present(x, y, rw, rh, rw, rh)
    base = np.array([[0.2, 0.2, 0.2]])
    gc = np.full((1, 3), S.godray_decompress(np.float64(np.float32(g))))
    for step, got in zip((0.0, 5.0), outs):
        blur = gc / 4.0 if step == 0.0 else np.zeros((1, 3))   # wW = 1 of wsum = 4, or wW = 0
        ref = S.composite_pixel(base, [base] * 4, np.array([20.0]), gc, blur, px, py, 4, np.zeros((1, 3)))
        assert np.abs(got - ref[0]).max() <= 2e-5, step


# ---------------------------------------------------------------- local TAA
def test_local_taa_formula(O):
    rng = np.random.default_rng(5)
    h, w = 24, 32
    cur = rng.uniform(-2.0, 4.0, (h, w, 4)).astype(np.float32)
    hist = rng.uniform(-1.0, 3.0, (h, w, 4)).astype(np.float32)
    cur[..., 3] = rng.choice([0.0, -0.5, 0.3, 1.0, 1.7], (h, w)).astype(np.float32)
    hist[..., 3] = rng.uniform(0.0, 1.0, (h, w)).astype(np.float32)
    # non-finite and out-of-range values in both inputs
    specials = np.array([np.nan, np.inf, -np.inf, 2e30, -5e30, 9e29], np.float32)
    for arr in (cur, hist):
        idx = rng.choice(h * w * 3, 60, replace=False)
        flat = arr[..., :3].reshape(-1)
        vals = rng.choice(specials, 60)
        view = arr[..., :3].copy().reshape(-1)
        view[idx] = vals
        arr[..., :3] = view.reshape(h, w, 3)
    fr = O.Frame(w, h)
    fr.local[...] = cur
    fr.local_hist[...] = hist
    O.local_taa(fr)
    ref = S.local_taa(cur.astype(np.float64), hist.astype(np.float64))
    got = fr.local_hist.astype(np.float64)
    assert np.array_equal(np.isnan(got), np.isnan(ref))
    ok = ~np.isnan(ref)
    assert np.allclose(got[ok], ref[ok], rtol=1e-6, atol=1e-6)
    # no valid sample: the confidence is kept bit for bit and the colour is the (cleaned) history
    none = ~((cur[..., 3] > 0) & ~S.is_bad_vec3(cur.astype(np.float64)))
    hist_ok = ~S.is_bad_vec3(hist.astype(np.float64))
    keep = none & hist_ok
    assert np.array_equal(fr.local_hist[keep][:, 3].view(np.uint32), hist[keep][:, 3].view(np.uint32))
    assert keep.sum() > 50 and (~none).sum() > 50


# ---------------------------------------------------------------- composite TAA
def test_composite_taa_passthrough_first_frames(O, vx):
    rng = np.random.default_rng(2)
    for fi in (0, 1):
        rw, rh = 24, 16
        cam = _static_cam(vx, rw, rh, frame_index=fi)
        fr = O.Frame(rw, rh)
        fr.pre_taa[...] = rng.uniform(0, 1, fr.pre_taa.shape).astype(np.float32)
        fr.out_hist[...] = rng.uniform(0, 1, fr.out_hist.shape).astype(np.float32)
        fr.depth[...] = rng.uniform(1, 100, fr.depth.shape).astype(np.float32)
        cur = fr.pre_taa.copy()
        O.composite_taa(cam, fr)
        assert np.array_equal(fr.output[..., :3], cur[..., :3]) and np.all(fr.output[..., 3] == 1.0)
        assert np.array_equal(fr.out_hist, fr.output)


def _taa_reference(vx, cam, rw, rh, pw, ph, depth, cur, hist):
    """mix(hist, cur, 0.45) at the history texel the pixel reprojects to, or
    cur outside the unit square.  Returns the reference for both texel
    choices on each axis (they differ only where the reprojection lands within
    1e-4 texel of a texel edge: there the float32 rounding of the camera
    tables decides) and the inside mask."""
    x, y = _grid(pw, ph)
    px, py = S.px_render_from_present(x, y, rw, rh, pw, ph)
    ix = np.clip(np.floor(px).astype(int), 0, rw - 1)
    iy = np.clip(np.floor(py).astype(int), 0, rh - 1)
    rd = S.ray_dir(cam, px, py)
    p = np.array(cam.cam_pos[:3], np.float64)[None, :] + rd * depth[iy, ix].astype(np.float64)[:, None]
    u, v = S.prev_uv(cam, p)
    inside = (u >= 0) & (v >= 0) & (u <= 1) & (v <= 1)
    tu, tv = u * pw, v * ph
    c = cur.reshape(-1, 4)[:, :3].astype(np.float64)
    refs = []
    for ex in (-1e-4, 1e-4):
        for ey in (-1e-4, 1e-4):
            hx = np.clip(np.floor(tu + ex).astype(int), 0, pw - 1)
            hy = np.clip(np.floor(tv + ey).astype(int), 0, ph - 1)
            refs.append(np.where(inside[:, None], S.mix(hist[hy, hx, :3].astype(np.float64), c, S.COMPOSITE_TAA_ALPHA), c))
    edge = np.minimum(np.abs(tu - np.round(tu)), np.abs(tv - np.round(tv))) < 1e-4
    border = (np.minimum(np.abs(u), np.abs(1 - u)) < 1e-7) | (np.minimum(np.abs(v), np.abs(1 - v)) < 1e-7)
    return refs, edge, inside, border


@pytest.mark.parametrize("render,present", [((32, 20), (32, 20)), ((16, 10), (32, 20))])
def test_composite_taa_static_and_moving_camera(O, vx, render, present):
    """The reference's ray table starts at the centre of pixel 0 (app.rs:454)
    and the shader adds the pixel centre again, so primary rays sit half a
    pixel right and down of what view_proj projects: at render = present a
    static camera reprojects every pixel exactly onto a texel corner, and
    either neighbouring texel may be taken there.  A moving camera lands
    inside texels."""
    (rw, rh), (pw, ph) = render, present
    rng = np.random.default_rng(3)
    for moving in (False, True):
        if moving:
            prev_cam = _cam(vx, rw, rh, pw, ph, frame_index=4, pos=(10.3, 20.1, 29.6), yaw=0.34, pitch=-0.13)
            cam = _cam(vx, rw, rh, pw, ph, frame_index=5, prev=np.array(prev_cam.view_proj, np.float32))
        else:
            cam = _static_cam(vx, rw, rh, pw, ph, frame_index=5)
        fr = O.Frame(rw, rh, pw, ph)
        fr.pre_taa[...] = rng.uniform(0, 1, fr.pre_taa.shape).astype(np.float32)
        fr.out_hist[...] = rng.uniform(0, 1, fr.out_hist.shape).astype(np.float32)
        fr.depth[...] = rng.uniform(2, 60, fr.depth.shape).astype(np.float32)
        cur, hist = fr.pre_taa.copy(), fr.out_hist.copy()
        O.composite_taa(cam, fr)
        refs, edge, inside, border = _taa_reference(vx, cam, rw, rh, pw, ph, fr.depth, cur, hist)
        ref_lo = refs[0]
        got = fr.output.reshape(-1, 4)[:, :3]
        cmp = ~border
        ok = np.zeros(len(got), bool)
        for r in refs:
            ok |= np.abs(got - r).max(-1) <= 1e-6
        assert ok[cmp].all(), (moving, np.argwhere(~ok & cmp)[:5])
        sel = ~edge & cmp
        assert not sel.any() or np.abs(got[sel] - ref_lo[sel]).max() <= 1e-6
        if not moving and (rw, rh) == (pw, ph):
            assert edge[inside].all()          # the half-pixel shift puts every reprojection on a texel corner
        else:
            assert (~edge).mean() > 0.95
        if moving:
            assert inside.mean() > 0.2


# ---------------------------------------------------------------- godray history
def _godray_frame(O, vx, rw, rh, depth, hist_lin, hist_depth, cam):
    w = O.World(np.zeros(1, vx.NODE_DTYPE), np.zeros(1, vx.ROPES_DTYPE), np.zeros(1, vx.META_DTYPE),
                np.full(1, 0xFFFFFFFF, np.uint32), np.zeros(16, np.uint32), np.zeros(2048, np.uint32), vx.Clipmap(),
                np.zeros((1, 1, 1), np.float32))
    fr = O.Frame(rw, rh)
    fr.depth[...] = depth
    comp = S.godray_compress(hist_lin).astype(np.float32)
    fr.godray_hist[..., :3] = comp
    fr.godray_hist[..., 3] = hist_depth
    O.godray(w, cam, fr)
    return fr.godray, comp.astype(np.float64)


def test_godray_fog_off_no_history_is_zero(O, vx):
    rw, rh = 24, 16
    cam = _static_cam(vx, rw, rh, fog=0.0, frame_index=2)
    out, _ = _godray_frame(O, vx, rw, rh, np.float32(40.0), np.zeros((rh, rw, 3)), np.float32(0.0), cam)
    assert np.all(out[..., :3] == 0.0) and np.all(out[..., 3] == 40.0)


def _godray_jitter(rw, rh, frame_index):
    """The sub-pixel ray jitter j_rd of compute_godray_pixel (godrays.wgsl:265-287),
    float32 arithmetic as the shader does it."""
    f32 = np.float32
    y, x = np.mgrid[0:rh, 0:rw]
    qx, qy = x.astype(f32), y.astype(f32)
    fi = f32(frame_index & 1023)
    fjx, fjy = fi, f32(fi * f32(1.371))
    jx = S.hash12((qx + fjx) + f32(1.7), (qy + fjy) + f32(9.2)) - 0.5
    jy = S.hash12((qx + f32(fjx * f32(1.13))) + f32(8.3), (qy + f32(fjy * f32(1.13))) + f32(2.1)) - 0.5
    return jx, jy


def test_godray_history_blend_closed_form(O, vx):
    """Fog off: the current term is zero, a constant depth makes the pixel
    'stable' (span 0), and a static camera reprojects pixel (x, y) to
    (x + 1 + 0.35 jx, y + 1 + 0.35 jy) in texel units: the shader's ray table
    puts rays half a pixel off the projection, plus the jitter.  So the output
    is compress(clamp(h, -0.015, 0.015) * mix(0.92, 0.03, react) * motion_ok)
    with h the decompressed history at that texel, react =
    smoothstep(0.10, 0.45, |h| / 0.051) and motion_ok = 1 - smoothstep(0.75,
    2.50, |(0.5 + 0.35 jx, 0.5 + 0.35 jy)|)."""
    rw, rh = 32, 24
    D = 37.5
    fi = 6
    cam = _static_cam(vx, rw, rh, fog=0.0, frame_index=fi)
    rng = np.random.default_rng(7)
    # per pixel either faint (|h| < 0.0052: no reaction) or up to 0.03 per channel (reactive, clamped)
    hist_lin = rng.uniform(0.0, 1.0, (rh, rw, 3)) * rng.choice([0.003, 0.03], (rh, rw, 1))
    out, comp = _godray_frame(O, vx, rw, rh, np.float32(D), hist_lin, np.float32(D), cam)
    h_all = S.godray_decompress(comp)
    jx, jy = _godray_jitter(rw, rh, fi)
    y, x = np.mgrid[0:rh, 0:rw]
    tx, ty = x + 1 + 0.35 * jx, y + 1 + 0.35 * jy
    sure = (np.abs(0.35 * jx) > 1e-4) & (np.abs(0.35 * jy) > 1e-4)
    inside = (tx <= rw) & (ty <= rh)
    hx = np.clip(np.floor(tx).astype(int), 0, rw - 1)
    hy = np.clip(np.floor(ty).astype(int), 0, rh - 1)
    h = h_all[hy, hx]
    motion_ok = 1.0 - S.smoothstep(0.75, 2.50, np.hypot(tx - (x + 0.5), ty - (y + 0.5)))
    react = S.smoothstep(0.10, 0.45, np.linalg.norm(h, axis=-1) / (0.05 + 1e-3))
    w = (S.mix(0.92, 0.03, react) * motion_ok * inside)[..., None]
    ref = S.godray_compress(np.maximum(np.clip(h, -0.015, 0.015) * w, 0.0))
    cmp = sure
    assert cmp.mean() > 0.99
    err = np.abs(out[..., :3] - ref)[cmp]
    assert err.max() <= 2e-6, err.max()
    assert np.all(out[..., 3] == np.float32(D))
    live = inside & sure
    assert (react[live] > 0.0).mean() > 0.3 and (react[live] == 0.0).mean() > 0.1
    assert (h[live] > 0.015).mean() > 0.2 and (motion_ok[live] < 1.0).mean() > 0.2   # every branch live

    # history from another depth: rel = |hist_depth - t| / t >= 0.2 fails the depth gate
    out2, _ = _godray_frame(O, vx, rw, rh, np.float32(D), hist_lin, np.float32(D * 1.25), cam)
    assert np.all(out2[..., :3] == 0.0)
    # partial gate: rel = 0.14 -> depth_ok = 1 - smoothstep(0.08, 0.20, 0.14)
    out3, _ = _godray_frame(O, vx, rw, rh, np.float32(D), hist_lin, np.float32(D * 1.14), cam)
    rel = abs(float(np.float32(D * 1.14)) - D) / D
    ok = 1.0 - S.smoothstep(0.08, 0.20, rel)
    ref3 = S.godray_compress(np.maximum(np.clip(h, -0.015, 0.015) * w * ok, 0.0))
    assert 0.1 < ok < 0.9 and np.abs(out3[..., :3] - ref3)[cmp].max() <= 2e-6


def test_godray_compress_roundtrip():
    c = np.linspace(0.0, 0.99, 1000)
    assert np.allclose(S.godray_compress(S.godray_decompress(c)), c, atol=1e-12)


# ---------------------------------------------------------------- sky
def _sky_world(O, vx):
    return O.World(np.zeros(1, vx.NODE_DTYPE), np.zeros(1, vx.ROPES_DTYPE), np.zeros(1, vx.META_DTYPE),
                   np.full(1, 0xFFFFFFFF, np.uint32), np.zeros(16, np.uint32), np.zeros(2048, np.uint32), vx.Clipmap(),
                   np.zeros((1, 1, 1), np.float32))


def _sun_pose_angles():
    yaw = math.atan2(S.SUN_DIR[0], S.SUN_DIR[2])
    return yaw, math.asin(S.SUN_DIR[1])


@pytest.mark.parametrize("height,pitch_off", [(300.0, 0.0), (300.0, -0.6), (60.0, -0.45)])
def test_sky_without_clouds_is_background_plus_sun(O, vx, height, pitch_off):
    """An empty world (no chunks, no clip levels): every primary pixel is
    sky_color(rd).  Above the cloud slab (y > 250 m), or below it for rays
    with rd.y <= 0.02 (the horizon fade zeroes the density), no cloud is
    marched, so the colour is sky_bg(rd) + sky_sun(rd) exactly."""
    rw, rh = 96, 64
    w = _sky_world(O, vx)
    yaw, pitch = _sun_pose_angles()
    cams = []
    for dyaw in (0.0, 0.25):
        pose = vx.make_pose((5.0, height, 7.0), yaw + dyaw, pitch + pitch_off, 40.0)
        cams.append(vx.camera_build(pose, (rw, rh), (rw, rh), 0, chunk_count=0))
    found_disc = False
    for cam, dyaw in zip(cams, (0.0, 0.25)):
        # the camera's forward vector sign convention is whatever camera_build uses: find the sun via the ray table
        fr = O.Frame(rw, rh)
        O.primary(w, cam, fr)
        x, y = _grid(rw, rh)
        px = x.astype(np.float32) + np.float32(0.5)
        py = y.astype(np.float32) + np.float32(0.5)
        rd = S.ray_dir(cam, px, py)
        ref = S.sky_bg(rd) + S.sky_sun(rd)
        mu = rd @ S.SUN_DIR
        edge = (mu > S.SUN_MU_OUTER - 2e-5) & (mu < S.SUN_MU_INNER + 2e-5)   # the disc edge's steep ramp
        clear = rd[:, 1] <= 0.02 if height < 250.0 else np.ones(len(rd), bool)
        cmp = clear & ~edge
        got = fr.color.reshape(-1, 4)
        assert np.all(fr.depth == S.FOG_MAX_DIST) and np.all(got[:, 3] == 1.0)
        err = np.abs(got[cmp, :3] - ref[cmp])
        assert cmp.sum() > rw * rh // 4
        assert err.max() <= 2e-5 * max(1.0, float(np.abs(ref[cmp]).max())), err.max()
        found_disc |= bool(np.any(mu[cmp] > S.SUN_MU_INNER))
    if height > 250.0 and pitch_off == 0.0:
        assert found_disc   # the disc itself is in view and compared
