"""Float64 statements of the reference's screen-space formulas, written from
the WGSL text (each function cites the lines it states), used to pin the
oracle on inputs where a pass reduces to these formulas: constant fields,
impulses, a static camera, fog switched off, rays that miss the cloud slab.

Test infrastructure: nothing in the product or the oracle imports this file.
The integer hashes are exact; everything else is evaluated in float64, so a
comparison with the oracle's float32 result needs a tolerance of a few
float32 ulps of the quantity compared.
"""
from __future__ import annotations

import numpy as np

LUM = np.array([0.2126, 0.7152, 0.0722])                 # composite.wgsl:6
SUN_DIR = np.array([0.61237244, 0.5, 0.61237244])       # common.wgsl:49
SUN_COLOR = np.array([1.0, 0.94, 0.72])                 # common.wgsl:50
SUN_INTENSITY = 5.0                                     # common.wgsl:51
SKY_EXPOSURE = 0.40                                     # common.wgsl:177
FOG_MAX_DIST = 120.0                                    # common.wgsl:84
LOCAL_TAA_ALPHA = 0.80                                  # common.wgsl:293
COMPOSITE_TAA_ALPHA = 0.45                              # common.wgsl:301
U32 = 0xFFFFFFFF


def clamp(x, a, b):
    return np.minimum(np.maximum(x, a), b)


def smoothstep(e0, e1, x):
    t = clamp((np.asarray(x, np.float64) - e0) / (e1 - e0), 0.0, 1.0)
    return t * t * (3.0 - 2.0 * t)


def mix(a, b, t):
    """WGSL mix(e1, e2, e3) = e1 * (1 - e3) + e2 * e3."""
    return a * (1.0 - t) + b * t


# ---- integer hashes (common.wgsl:595-619) --------------------------------
def hash_u32(v):
    v = np.asarray(v, np.uint64) & U32
    v ^= v >> np.uint64(16)
    v = (v * np.uint64(0x7FEB352D)) & U32
    v ^= v >> np.uint64(15)
    v = (v * np.uint64(0x846CA68B)) & U32
    v ^= v >> np.uint64(16)
    return v


def hash2_u32(x, y):
    x = np.asarray(x, np.uint64) & U32
    y = np.asarray(y, np.uint64) & U32
    return hash_u32(((x * np.uint64(0x8DA6B343)) & U32) ^ ((y * np.uint64(0xD8163841)) & U32))


def hash12(px, py):
    """hash12 of float32 coordinates: floor, i32, bitcast u32, hash, f32(h) * 2^-32."""
    ix = np.floor(np.asarray(px, np.float32)).astype(np.int64) & U32
    iy = np.floor(np.asarray(py, np.float32)).astype(np.int64) & U32
    h = hash2_u32(ix, iy).astype(np.uint32)
    return (h.astype(np.float32) * np.float32(2.0 ** -32)).astype(np.float64)


# ---- present <-> render mapping (common.wgsl:664-690) ---------------------
def px_render_from_present(x, y, rw, rh, pw, ph):
    """Render-space pixel position (float32, as the shader computes it) of the
    centre of present pixel (x, y)."""
    px = (np.asarray(x, np.float32) + np.float32(0.5)) / np.float32(pw) * np.float32(rw)
    py = (np.asarray(y, np.float32) + np.float32(0.5)) / np.float32(ph) * np.float32(rh)
    return px.astype(np.float32), py.astype(np.float32)


# ---- tonemap and composite (composite.wgsl:5-204, ray_main.wgsl:537-573) ---
def bright_extract_hue(x, thresh):
    l = x @ LUM
    w = np.maximum(l - thresh, 0.0) / np.maximum(l, 1e-4)
    return x * w[..., None]


def hable(x):
    A, B, C, D, E, F = 0.22, 0.30, 0.10, 0.20, 0.01, 0.30
    num = x * (A * x + C * B) + D * E
    den = x * (A * x + B) + D * F
    return num / np.maximum(den, 1e-6) - E / F


def hable_white(x, white_point=6.0):
    return hable(x) / np.maximum(hable(np.float64(white_point)), 1e-6)


def godray_decompress(c):   # godrays.wgsl:222-227
    return 0.25 * c / np.maximum(1.0 - c, 1e-4)


def godray_compress(lin):   # godrays.wgsl:229-233
    return lin / (lin + 0.25)


def composite_pixel(base, bloom_taps, d0, g_c, g_blur, px, py, frame_index, local):
    """composite_pixel_mapped + the local add of main_composite.
    base (N,3): colour at the mapped render pixel; bloom_taps: four (N,3)
    colour taps at +-2 pixels; d0 (N,): depth at the mapped full pixel;
    g_c / g_blur (N,3): decompressed centre godray tap and the depth-weighted
    blur of the four neighbour taps; (px, py): float32 render position;
    local (N,3): the local-light history sample."""
    far = smoothstep(40.0, 80.0, d0)[:, None]                          # GODRAY_FADE_NEAR/FAR
    god = np.maximum(g_c + 0.1 * (g_c - g_blur), 0.0)                   # COMPOSITE_SHARPEN
    god = np.maximum(god - 0.004, 0.0)                                  # GODRAY_BLACK_LEVEL
    god = god / (god + 0.25)                                            # GODRAY_KNEE_COMPOSITE
    scale = 6.5 * mix(1.0, 0.25, far)                                   # GODRAY_COMPOSITE_SCALE
    sun_hue = np.maximum(SUN_COLOR, 1e-4)
    sun_hue = sun_hue / np.sqrt(sun_hue @ sun_hue)
    god_sun = sun_hue * np.maximum(god @ LUM, 0.0)[:, None]
    god = mix(god, god_sun, 0.55 * far)
    god = god * mix(np.ones(3), np.array([1.08, 1.03, 0.92]), far)
    hdr = np.maximum(base + scale * god, 0.0)
    k_eff = 0.12 * mix(1.0, 0.0, far)
    bloom = bright_extract_hue(hdr, 1.4)
    for tap in bloom_taps:
        bloom = bloom + bright_extract_hue(np.maximum(tap, 0.0), 1.4)
    bloom = bloom / 5.0
    hdr = hdr + k_eff * np.minimum(bloom, 0.35 * np.maximum(hdr, 0.0))
    l_hdr = np.maximum(hdr @ LUM, 1e-6)
    sat = 1.0 + 0.55 * smoothstep(30.0, 100.0, d0)
    sat = mix(sat, 1.0, 0.55 * smoothstep(1.6, 6.0, l_hdr))
    hdr = mix(l_hdr[:, None], hdr, sat[:, None])
    hdr = hdr * 1.10 * np.array([1.03, 1.0, 0.97])
    hdr = np.maximum((hdr - 0.18) * 1.05 + 0.18, 0.0)
    c = clamp(hable_white(hdr * 0.15), 0.0, 1.0)                        # POST_EXPOSURE
    c = c ** 0.98
    fi = np.float32(frame_index & 255)
    n0 = hash12(px + fi, py) - 0.5
    n1 = hash12((px * np.float32(0.73)).astype(np.float32), (py * np.float32(0.73)).astype(np.float32) + fi) - 0.5
    c = c + ((0.6 * n0 + 0.4 * n1) / 1536.0)[:, None]
    return clamp(c, 0.0, 1.0) ** (1.0 / 2.2) + local


# ---- temporal passes (taa.wgsl:27-110) ----------------------------------
def is_bad_vec3(v):   # common.wgsl:692-706
    return np.any(np.isnan(v) | (np.abs(v) > 1e30), axis=-1)


def local_taa(cur4, hist4):
    hist_ok = ~is_bad_vec3(hist4[..., :3])
    hist = np.where(hist_ok[..., None], hist4[..., :3], 0.0)
    conf0 = np.where(hist_ok, hist4[..., 3], 0.0)
    cur, w_in = cur4[..., :3], cur4[..., 3]
    has = (w_in > 0.0) & ~is_bad_vec3(cur)
    k = np.where(has, LOCAL_TAA_ALPHA * clamp(w_in, 0.0, 1.0), 0.0)
    with np.errstate(invalid="ignore"):
        out = mix(hist, cur, k[..., None])
    conf1 = np.where(has, np.minimum(1.0, conf0 + k), conf0)
    return np.concatenate([out, conf1[..., None]], axis=-1)


# ---- camera helpers (common.wgsl:456-471) --------------------------------
def ray_dir(cam, px, py):
    f = lambda a: np.array(a[:3], np.float64)
    d = f(cam.ray00)[None, :] + np.asarray(px, np.float64)[:, None] * f(cam.ray_dx) + \
        np.asarray(py, np.float64)[:, None] * f(cam.ray_dy)
    return d / np.linalg.norm(d, axis=-1, keepdims=True)


def prev_uv(cam, p):
    m = np.array(cam.prev_view_proj, np.float64).reshape(4, 4)  # column-major: m[col, row]
    ph = np.concatenate([p, np.ones((len(p), 1))], axis=-1)
    clip = ph @ m
    w = np.maximum(clip[:, 3], 1e-6)
    u = clip[:, 0] / w * 0.5 + 0.5
    v = clip[:, 1] / w * 0.5 + 0.5
    return u, 1.0 - v


# ---- sky (sky.wgsl:5-41) ------------------------------------------------
def sky_bg(rd):
    y = rd[:, 1]
    col = mix(np.array([0.05, 0.08, 0.12]), np.array([0.55, 0.75, 0.95]), smoothstep(-0.10, 0.85, y)[:, None])
    col = col + (0.12 * (1.0 - smoothstep(-0.20, 0.18, np.abs(y))))[:, None] * np.array([0.80, 0.75, 0.70])
    col = col + (0.05 * smoothstep(0.15, 1.00, y))[:, None] * np.array([0.20, 0.35, 0.60])
    return col * SKY_EXPOSURE


SUN_MU_INNER = np.cos(0.009)
SUN_MU_OUTER = np.cos(0.009 + 0.004)


def sky_sun(rd):
    mu = clamp(rd @ SUN_DIR, -1.0, 1.0)
    disc = smoothstep(SUN_MU_OUTER, SUN_MU_INNER, mu)
    glow = smoothstep(0.70, 0.98, mu)
    halo = 0.03 * np.exp(-2500.0 * np.maximum(0.0, 1.0 - mu))
    return SUN_COLOR * SUN_INTENSITY * (disc + halo + 0.04 * glow)[:, None]


# ---- value noise and fbm (common.wgsl:614-662) ---------------------------
def _corner(ix, iy):
    """hash12 at an integer lattice point: the u32 hash rounded to float32, times 2^-32."""
    h = hash2_u32(np.asarray(ix, np.int64) & U32, np.asarray(iy, np.int64) & U32).astype(np.uint32)
    return h.astype(np.float32).astype(np.float64) * 2.0 ** -32


def value_noise(x, y):
    ix, iy = np.floor(x), np.floor(y)
    fx, fy = x - ix, y - iy
    ix, iy = ix.astype(np.int64), iy.astype(np.int64)
    a, b = _corner(ix, iy), _corner(ix + 1, iy)
    c, d = _corner(ix, iy + 1), _corner(ix + 1, iy + 1)
    ux, uy = fx * fx * (3.0 - 2.0 * fx), fy * fy * (3.0 - 2.0 * fy)
    return mix(mix(a, b, ux), mix(c, d, ux), uy)


def fbm(x, y):
    """Five octaves; each step maps p -> R p * 2 + (17, 9) with the WGSL
    mat2x2(0.8, -0.6, 0.6, 0.8), whose columns are (0.8, -0.6) and (0.6, 0.8)."""
    s, amp = 0.0, 0.5
    for _ in range(5):
        s = s + amp * value_noise(x, y)
        x, y = (0.8 * x + 0.6 * y) * 2.0 + 17.0, (-0.6 * x + 0.8 * y) * 2.0 + 9.0
        amp *= 0.5
    return s


# ---- phase and clouds (phase.wgsl, clouds.wgsl, common.wgsl:124-175) -----
INV_4PI = 0.0795774715
CLOUD_BASE_H, CLOUD_TOP_H = 170.0, 250.0
CLOUD_UV_SCALE = 0.0016
CLOUD_WIND = np.array([0.020, 0.012])
CLOUD_DENSITY = 0.058
CLOUD_SHADOW_ABSORB = 6.0


def phase_blended(costh):
    g = 0.15
    mie = INV_4PI * (1.0 - g * g) / np.maximum((1.0 + g * g - 2.0 * g * costh) ** 1.5, 1e-3)
    return mix(INV_4PI, mie, 0.20)


def cloud_height01(y):
    t = clamp((y - CLOUD_BASE_H) / (CLOUD_TOP_H - CLOUD_BASE_H), 0.0, 1.0)
    return smoothstep(0.0, 0.12, t) * (1.0 - smoothstep(0.78, 1.0, t))


def cloud_coverage_at_xz(x, z, time_s):
    u, v = x * CLOUD_UV_SCALE + CLOUD_WIND[0] * time_s, z * CLOUD_UV_SCALE + CLOUD_WIND[1] * time_s
    field = 0.65 * fbm(u, v) + 0.35 * fbm(u * 2.3 + 13.2, v * 2.3 + 7.1)
    return smoothstep(0.40, 0.48, field)


def cloud_density(p, time_s):
    """p (N, 3) -> (N,); zero outside the slab, where the height profile is
    zero and where the coverage is at most 0.001."""
    x, y, z = p[:, 0], p[:, 1], p[:, 2]
    out = np.zeros(len(p))
    h = cloud_height01(y)
    live = (y > CLOUD_BASE_H) & (y < CLOUD_TOP_H) & (h > 0.0)
    cov = np.zeros(len(p))
    cov[live] = cloud_coverage_at_xz(x[live], z[live], time_s)
    live &= cov > 0.001
    if not live.any():
        return out
    x, y, z, cov, h = x[live], y[live], z[live], cov[live], h[live]
    u = x * CLOUD_UV_SCALE + CLOUD_WIND[0] * time_s + 0.003 * y
    v = z * CLOUD_UV_SCALE + CLOUD_WIND[1] * time_s + 0.003 * y
    base = fbm(u * 1.2, v * 1.2)
    det = fbm(u * 3.7 + 0.11 * y, v * 3.7 - 0.07 * y)
    d = smoothstep(0.45, 0.95, mix(base, det, 0.35)) ** 1.8
    out[live] = clamp(d * cov * h, 0.0, 1.0)
    return out


def cloud_light_transmittance(p, time_s):
    tmax = (CLOUD_TOP_H - p[:, 1]) / SUN_DIR[1]
    dt = tmax / 4.0
    od = np.zeros(len(p))
    for i in range(4):
        od += cloud_density(p + SUN_DIR[None, :] * ((i + 0.5) * dt)[:, None], time_s) * dt
    return np.where(tmax <= 0.0, 1.0, np.exp(-CLOUD_DENSITY * od * CLOUD_SHADOW_ABSORB))


def _slab_interval(p, d):
    t0 = (CLOUD_BASE_H - p[:, 1]) / d[1]
    t1 = (CLOUD_TOP_H - p[:, 1]) / d[1]
    return np.minimum(t0, t1), np.maximum(t0, t1)


def cloud_sun_transmittance_fast(p, time_s):
    """2-step cloud optical depth towards the sun (clouds.wgsl:123-154)."""
    ta, tb = _slab_interval(p, SUN_DIR)
    t0, t1 = np.maximum(ta, 0.0), np.maximum(tb, 0.0)
    dt = (t1 - t0) / 2.0
    od = np.zeros(len(p))
    for i in range(2):
        od += cloud_density(p + SUN_DIR[None, :] * (t0 + (i + 0.5) * dt)[:, None], time_s) * dt
    return np.where(t1 <= t0, 1.0, np.exp(-CLOUD_DENSITY * od * CLOUD_SHADOW_ABSORB))


def cloud_density_godray(p, time_s):
    x, y, z = p[:, 0], p[:, 1], p[:, 2]
    h = cloud_height01(y)
    u, v = x * CLOUD_UV_SCALE + CLOUD_WIND[0] * time_s, z * CLOUD_UV_SCALE + CLOUD_WIND[1] * time_s
    field = 0.70 * value_noise(u, v) + 0.30 * value_noise(u * 2.0 + 17.0, v * 2.0 + 9.0)
    d = smoothstep(0.40, 0.48, field) * h
    return np.where((y <= CLOUD_BASE_H) | (y >= CLOUD_TOP_H) | (h <= 0.0), 0.0, d)


def cloud_sun_transmittance_godray(p, time_s):
    """One density sample at the midpoint of the sun ray's slab segment (clouds.wgsl:176-215)."""
    ta, tb = _slab_interval(p, SUN_DIR)
    t0 = np.maximum(ta, 0.0)
    pm = p + SUN_DIR[None, :] * (t0 + 0.5 * (tb - t0))[:, None]
    tr = np.exp(-CLOUD_DENSITY * cloud_density_godray(pm, time_s) * (tb - t0) * CLOUD_SHADOW_ABSORB)
    return np.where((tb <= 0.0) | (tb <= t0), 1.0, tr)


def sky_color(rd, ro, time_s, margin=None):
    """sky_color (sky.wgsl:43-137) with the 8-step cloud march.  If `margin` is
    a dict, margin['near'] flags rays where a float32 evaluation could take
    the other side of the march's branches (density against 1e-4,
    transmittance against 0.02)."""
    n = len(rd)
    bg, sun = sky_bg(rd), sky_sun(rd)
    T = np.ones(n)
    acc = np.zeros((n, 3))
    near = np.zeros(n, bool)
    up = rd[:, 1] > 0.01
    with np.errstate(divide="ignore", invalid="ignore"):
        t0 = (CLOUD_BASE_H - ro[1]) / rd[:, 1]
        t1 = (CLOUD_TOP_H - ro[1]) / rd[:, 1]
    te, tx = np.maximum(np.minimum(t0, t1), 0.0), np.maximum(t0, t1)
    live = up & (tx > te)
    horizon = clamp((rd[:, 1] - 0.02) / 0.25, 0.0, 1.0)
    dt = np.where(live, (tx - te) / 8.0, 0.0)
    mu = rd @ SUN_DIR
    phase = phase_blended(mu)
    silver = clamp(mu, 0.0, 1.0) ** 8.0 * 0.6
    ccol = mix(np.array([0.72, 0.74, 0.76])[None, :], 1.0, silver[:, None])
    going = live.copy()
    for i in range(8):
        idx = np.nonzero(going)[0]
        if len(idx) == 0:
            break
        ti = te[idx] + (i + 0.5) * dt[idx]
        p = ro[None, :] + rd[idx] * ti[:, None]
        dens = cloud_density(p, time_s) * horizon[idx]
        near[idx] |= np.abs(dens - 1e-4) < 1e-6
        act = dens > 1e-4
        j = idx[act]
        if len(j):
            tl = cloud_light_transmittance(p[act], time_s)
            a = 1.0 - np.exp(-(CLOUD_DENSITY * dens[act] * dt[j]))
            scatter = ccol[j] * (SUN_COLOR * SUN_INTENSITY)[None, :] * (phase[j] * tl)[:, None]
            acc[j] += (T[j] * a)[:, None] * scatter
            T[j] *= 1.0 - a
            near[j] |= np.abs(T[j] - 0.02) < 1e-5
            going[j[T[j] < 0.02]] = False
    col = bg * T[:, None] + acc
    t_sun = np.where(up, clamp(T, 0.0, 1.0) ** 8.0, 1.0)
    if margin is not None:
        margin["near"] = near
        margin["T"] = T
    return col + sun * t_sun[:, None]


# ---- fog (fog.wgsl) -------------------------------------------------------
FOG_HEIGHT_FALLOFF = 0.18
FOG_COLOR_GROUND = np.array([0.40, 0.42, 0.45])


def fog_optical_depth(base, ro, rd, t):
    k, y0, dy = FOG_HEIGHT_FALLOFF, ro[1], rd[:, 1]
    with np.errstate(divide="ignore", invalid="ignore"):
        gen = base * (np.exp(-k * y0) - np.exp(-k * (y0 + dy * t))) / (k * dy)
    return np.where(np.abs(dy) < 1e-4, base * np.exp(-k * y0) * t, gen) if base > 0.0 else np.zeros(len(rd))


def apply_fog(surface, ro, rd, t, sky, fog_param):
    T = np.exp(-np.maximum(fog_optical_depth(max(fog_param * 0.02, 0.0), ro, rd, t), 0.0))
    up = clamp(rd[:, 1] * 0.5 + 0.5, 0.0, 1.0)
    fogc = mix(FOG_COLOR_GROUND[None, :], np.minimum(sky, 0.45), (0.10 * up)[:, None])
    ins = fogc + (0.35 * clamp(rd @ SUN_DIR, 0.0, 1.0) ** 8.0)[:, None] * SUN_COLOR[None, :]
    return surface * T[:, None] + ins * (1.0 - T)[:, None]


# ---- godray march (godrays.wgsl:22-216, 241-356) -------------------------
def exp_neg_fast(x):
    x = np.maximum(x, 0.0)
    return 1.0 / (1.0 + x + 0.48 * x * x)


def godray_pixel_inputs(cam, depth):
    """The per-pixel inputs of compute_godray_pixel that are fixed by float32
    hashing: j4, j_phase, the quantized t_end, the jittered ray direction and
    the step count.  depth (rh, rw) float32 is the centre tap; arithmetic that
    feeds a floor() is done in float32 exactly as the shader orders it."""
    f = np.float32
    rh, rw = depth.shape
    y, x = np.mgrid[0:rh, 0:rw]
    qx, qy = x.astype(f).ravel(), y.astype(f).ravel()
    fi = f(cam.frame_index & 1023)
    fjx, fjy = fi, f(fi * f(1.371))
    j4 = 0.20 * (hash12(qx * f(1.31) + fjx, qy * f(1.31) + fjy) - 0.5)
    j_phase = hash12((qx * f(0.91) + fjx * f(0.73)) + f(31.7), (qy * f(0.91) + f(fjy * f(0.73))) + f(12.3)) - 0.5
    dq = f(hash12((qx * f(1.37) + fjx * f(1.19)) + f(9.2), (qy * f(1.37) + f(fjy * f(1.19))) + f(1.1)) - 0.5) * f(0.03)
    jx = hash12((qx + fjx) + f(1.7), (qy + fjy) + f(9.2)) - 0.5
    jy = hash12((qx + f(fjx * f(1.13))) + f(8.3), (qy + f(fjy * f(1.13))) + f(2.1)) - 0.5
    tc = depth.ravel().astype(f)
    q = (tc + dq) / f(0.03)
    t_end = np.minimum(np.floor(q) * f(0.03), f(100.0)).astype(np.float64)
    near = np.abs(q - np.round(q)) < 1e-4                       # the quantizer's floor could fall either way
    raw = np.ceil((t_end.astype(f) * f(1.25)).astype(np.float64))
    s = np.clip(raw, 8, 24).astype(np.int64)
    steps = np.minimum(((s + 3) // 4) * 4, 24)
    px = qx + f(0.5) + f(0.35) * jx.astype(f)
    py = qy + f(0.5) + f(0.35) * jy.astype(f)
    return {"j4": j4, "j_phase": j_phase, "t_end": t_end, "steps": steps, "rd": ray_dir(cam, px, py), "near": near}


def godray_integrate(ro, rd, t_end, j, j_phase, steps, fog_param, time_s, ts_geom=None):
    """godray_integrate_1 for N rays (float64).  ts_geom(p, idx) -> the
    geometry-only sun transmittance at sample points p of rays idx (default: 1,
    an empty world).  Returns (rgb (N,3), near (N,)): near flags rays where a
    float32 evaluation could take the other side of the march's cut-offs."""
    n = len(rd)
    base = max(fog_param * 1.0, 0.0)
    out = np.zeros((n, 3))
    near = np.zeros(n, bool)
    if base <= 0.0:
        return out, near
    mu = rd @ SUN_DIR
    phase = phase_blended(mu)
    view_gate = 0.10 + 0.90 * smoothstep(0.0, 0.20, mu)
    dt = t_end / steps
    a_ts, a_shaft = 1.0 - np.exp(-dt * 3.0), 1.0 - np.exp(-dt * 5.0)
    k, y0, dy = FOG_HEIGHT_FALLOFF, ro[1], rd[:, 1]
    fr = clamp(0.5 + j_phase, 0.0, 1.0)
    horiz = np.abs(dy) < 1e-4
    near |= np.abs(np.abs(dy) - 1e-4) < 1e-8
    dens_const = base * np.exp(-k * y0)
    ts_h, te_h = exp_neg_fast(dens_const * dt * fr), exp_neg_fast(dens_const * dt * (1.0 - fr))
    with np.errstate(divide="ignore", invalid="ignore"):
        inv_kdy = np.where(horiz, 0.0, 1.0 / (k * dy))
    B = np.full(n, np.exp(-k * y0))
    r_f, r_rem = np.exp(-k * dy * dt * fr), np.exp(-k * dy * dt * (1.0 - fr))
    tv = np.ones(n)
    ts = np.ones(n)
    sh = np.zeros(n)
    ts_g = np.ones(n)
    acc = np.zeros((n, 3))
    going = t_end > 0.0
    for i in range(24):
        going &= i < steps
        ti = (i + 0.5 + j_phase) * dt
        cut = ti + j * dt
        near |= going & (np.abs(cut - t_end) < 1e-5 * np.maximum(t_end, 1.0))
        going &= ~(cut > t_end)
        idx = np.nonzero(going)[0]
        if len(idx) == 0:
            break
        p = ro[None, :] + rd[idx] * ti[idx, None]
        bs = B[idx] * r_f[idx]
        tv[idx] *= np.where(horiz[idx], ts_h[idx], exp_neg_fast(np.maximum(base * (B[idx] - bs) * inv_kdy[idx], 0.0)))
        near[idx] |= np.abs(tv[idx] - 0.02) < 1e-6
        stop = tv[idx] < 0.02
        going[idx[stop]] = False
        idx, p, bs = idx[~stop], p[~stop], bs[~stop]
        if len(idx) == 0:
            break
        if i % 4 == 0:
            ts_g[idx] = 1.0 if ts_geom is None else ts_geom(p, idx)
        tc = cloud_sun_transmittance_godray(p, time_s)
        tc_vol = mix(1.0, tc, 0.50)
        y1 = clamp(ts_g[idx], 0.0, 1.0)
        ts_soft = mix(y1, np.sqrt(y1), 0.5)
        haze_ramp = 1.0 - np.exp(-ti[idx] / 18.0)
        dens = base * np.maximum(np.exp(-0.04 * p[:, 1]), 0.35)
        common = (SUN_COLOR * SUN_INTENSITY)[None, :] * (dens * dt[idx] * tv[idx] * phase[idx] * tc_vol * 0.70)[:, None]
        ts_prev = ts[idx]
        ts[idx] = mix(ts_prev, ts_soft, a_ts[idx])
        d_ts = np.maximum(0.0, ts_prev - ts_soft)
        shaft = smoothstep(0.004, 0.035, d_ts) * (1.0 - ts_soft)
        s1 = clamp(shaft, 0.0, 1.0)
        shaft = mix(np.sqrt(s1), s1, 0.20)
        sh[idx] = mix(sh[idx], shaft, a_shaft[idx])
        shaft = sh[idx]
        edge_boost = 1.0 + 2.5 * shaft
        gate = smoothstep(0.35, 0.80, ts[idx])
        haze = 0.02 * haze_ramp * ts[idx] * ts[idx]
        wgt = clamp(haze + (1.0 - haze) * (shaft * gate), 0.0, 1.0)
        acc[idx] += common * (view_gate[idx] * edge_boost * ts[idx] * wgt)[:, None]
        be = bs * r_rem[idx]
        tv[idx] *= np.where(horiz[idx], te_h[idx], exp_neg_fast(np.maximum(base * (bs - be) * inv_kdy[idx], 0.0)))
        B[idx] = np.where(horiz[idx], B[idx], be)
    g = acc * 8.0
    return g / (g + 0.35), near


# ---- voxel surface shading (shading.wgsl:7-52,210-362; ray_main.wgsl:428-470) --
MAT_COLOR = {2: (0.45, 0.30, 0.15), 3: (0.50, 0.50, 0.55), 4: (0.38, 0.26, 0.14)}   # dirt, stone, wood
MAT_ROUGH = {2: 0.90, 3: 0.45, 4: 0.70}
MAT_F0 = {2: 0.02, 3: 0.04, 4: 0.03}


def hash31_cells(cx, cy, cz):
    """hash31 at integer cells: the u32 hash3 rounded to float32, times 2^-32."""
    m = lambda a: (np.asarray(a, np.int64) & U32).astype(np.uint64)
    h = hash_u32(((m(cx) * np.uint64(0x8DA6B343)) & U32) ^ ((m(cy) * np.uint64(0xD8163841)) & U32)
                 ^ ((m(cz) * np.uint64(0xCB1AB31F)) & U32)).astype(np.uint32)
    return h.astype(np.float32).astype(np.float64) * 2.0 ** -32


def material_albedo(mat, cells, t):
    """base_albedo (shading.wgsl:196-202) for dirt, stone or wood: the material
    colour, varied within FAR_SHADING_DIST by ALBEDO_VAR_GAIN (3.5) times the
    hash of the hit's 0.05 m cell (cells: (N, 3) integer cell indices)."""
    base = np.array(MAT_COLOR[mat], np.float64)[None, :].repeat(len(t), 0)
    v = 3.5 * (hash31_cells(cells[:, 0], cells[:, 1], cells[:, 2]) - 0.5) * 2.0
    c = base.copy()
    if mat == 2:
        c = (c + np.stack([0.04 * v, 0.02 * v, 0.01 * v], -1)) * (1.0 + 0.08 * v)[:, None]
    elif mat == 3:
        c = c * (1.0 + 0.10 * v)[:, None]
    elif mat == 4:
        c = (c + np.stack([0.05 * v, 0.02 * v, 0.0 * v], -1)) * (1.0 + 0.07 * v)[:, None]
    c = clamp(c, 0.0, 1.5)
    return np.where((t <= 64.0)[:, None], c, base)


def shade_voxel(ro, rd, t, n, mat, albedo, ao, sun_shadow, time_s):
    """shade_hit (shading.wgsl:263-362) for a dirt, stone or wood voxel face:
    no grass, dapple or emission terms.  n (N, 3) the face normal, ao (N,) the
    macro AO term (used within VOXEL_AO_MAX_DIST), sun_shadow (N,) the
    history-blended geometric shadow."""
    hp = ro[None, :] + rd * t[:, None]
    tc = cloud_sun_transmittance_fast(hp + n * 0.075, time_s)
    sun_vis = tc * clamp(sun_shadow, 0.0, 1.0)
    diff = np.maximum(n @ SUN_DIR, 0.0)
    ao = np.where(t <= 40.0, ao, 1.0)
    sky_up = sky_bg(np.array([[0.0, 1.0, 0.0]]))[0]
    amb_col = mix(FOG_COLOR_GROUND[None, :], sky_up[None, :], clamp(n[:, 1] * 0.5 + 0.5, 0.0, 1.0)[:, None])
    ambient = amb_col * 0.10 * clamp(ao, 0.0, 1.0)[:, None]
    if mat == 3:
        ambient = ambient * np.array([0.92, 0.95, 1.05])
    v = -rd / np.linalg.norm(rd, axis=-1, keepdims=True)
    h = v + SUN_DIR[None, :]
    h = h / np.linalg.norm(h, axis=-1, keepdims=True)
    ndv = np.maximum(np.sum(n * v, -1), 0.0)
    ndh = np.maximum(np.sum(n * h, -1), 0.0)
    shin = mix(8.0, 96.0, 1.0 - MAT_ROUGH[mat])
    f0 = MAT_F0[mat]
    fres = f0 + (1.0 - f0) * (1.0 - clamp(ndv, 0.0, 1.0)) ** 5.0
    spec = (SUN_COLOR * SUN_INTENSITY)[None, :] * (ndh ** shin * fres * sun_vis)[:, None]
    spec = np.where((t <= 64.0)[:, None], spec, 0.0)
    direct = (SUN_COLOR * SUN_INTENSITY)[None, :] * ((0.35 * diff + 0.65 * diff * diff) * sun_vis)[:, None]
    return albedo * (ambient + direct) + 0.20 * spec


# ---- heightfield hits: the flat-clipmap trace and the grass-top shading -----
def clip_trace_flat(ro, rd, height, cell, t_max=FOG_MAX_DIST, eps=1e-5):
    """clip_trace_heightfield (clipmap.wgsl:156-245) over a single clip level
    whose texels all hold `height` (nearest and bilinear samples are then the
    constant, the 2-tap normal is +y).  The march's vertical step lands on the
    plane itself (dt = |s| / -rd.y), where float32 rounding decides the sign
    of the next signed height; wherever a signed height is within `eps` of
    zero both outcomes are followed.  Returns, per ray, the list of hit
    distances the trace can produce (empty: no hit)."""
    vy = np.maximum(-rd[:, 1], 0.12)
    vh = np.maximum(np.hypot(rd[:, 0], rd[:, 2]), 1e-4)
    out = []
    for k in range(len(rd)):
        if rd[k, 1] >= -1e-4:
            out.append([])
            continue
        found = set()

        def bisect(a, b, depth):
            if depth == 3:
                found.add(0.5 * (a + b))
                return
            m = 0.5 * (a + b)
            sm = ro[1] + rd[k, 1] * m - height
            if sm > -eps:
                bisect(m, b, depth + 1)
            if sm < eps:
                bisect(a, m, depth + 1)

        def march(t, t_prev, s_prev, i):
            while i < 48 and t <= t_max:
                s = ro[1] + rd[k, 1] * t - height
                if s_prev > 0.0 and s < eps:
                    bisect(t_prev, t, 0)
                    if s <= -eps:
                        return
                    s = max(s, 0.0) + 1e-30   # the other outcome: a tiny positive height, march on
                s_prev, t_prev = s, t
                t = t + min(max(min(abs(s) / vy[k], 6.0 * cell / vh[k]), 0.6), 48.0)
                i += 1

        march(0.0, 0.0, ro[1] - height, 0)
        out.append(sorted(found))
    return out


def grass_blade_slivers(px, pz, freq, detail, flags):
    """grass_blade_slivers (shading.wgsl:64-94) where the wind field is zero
    (hit heights at or below WIND_RAMP_Y0 = 2 m): the blade direction is
    normalize((0.6, 0.2)).  flags['near'] marks points on a sliver cell edge."""
    d = np.array([0.6, 0.2]) / np.hypot(0.6, 0.2)
    u0 = (px * d[0] + pz * d[1]) * freq
    v0 = (px * -d[1] + pz * d[0]) * freq
    cx, cy = np.floor(u0), np.floor(v0)
    fx, fy = u0 - cx, v0 - cy
    flags["near"] |= (np.minimum(fx, 1.0 - fx) < 2e-3) | (np.minimum(fy, 1.0 - fy) < 2e-3)
    cx, cy = cx.astype(np.int64), cy.astype(np.int64)
    h0 = hash31_cells(cx, cy, 11)
    h1 = hash31_cells(cx + 3, cy + 7, 5)
    ang = (h0 - 0.5) * 0.35
    prx = fx * np.cos(ang) - fy * np.sin(ang)
    pry = fx * np.sin(ang) + fy * np.cos(ang)
    center = 0.5 + (h1 - 0.5) * 0.6
    u = clamp(prx, 0.0, 1.0)
    width = mix(0.065, 0.012, u ** 1.8) * mix(0.75, 1.25, h0)
    aa = mix(1.6, 0.6, detail) / freq
    sliver = smoothstep(width + aa, width - aa, np.abs(pry - center))
    return sliver * smoothstep(0.0, 0.15, u) * smoothstep(1.0, 0.75, u)


def grass_top_albedo(base, hp, t, flags):
    """grass_top_albedo (shading.wgsl:107-135) with the zero-wind blade field."""
    tip = clamp(hp[:, 1] / 0.1 - np.floor(hp[:, 1] / 0.1), 0.0, 1.0)
    fr = hp[:, 1] / 0.1 - np.floor(hp[:, 1] / 0.1)
    flags["near"] |= np.minimum(fr, 1.0 - fr) < 1e-4
    detail = 1.0 - smoothstep(18.0, 55.0, t)
    low = grass_blade_slivers(hp[:, 0], hp[:, 2], 9.0, detail * 0.6, flags)
    high = grass_blade_slivers(hp[:, 0] + 1.7, hp[:, 2] + 2.3, 18.0, detail, flags)
    blade = mix(low, 0.5 * (low + high), mix(0.15, 1.0, detail))
    patches = fbm(hp[:, 0] * 0.08, hp[:, 2] * 0.08)
    clumps = fbm(hp[:, 0] * 0.35 + 12.3, hp[:, 2] * 0.35 + 7.1)
    cov = (blade * smoothstep(0.25, 0.75, patches) * smoothstep(0.20, 0.80, clumps))[:, None]
    grad = smoothstep(0.0, 1.0, tip)[:, None]
    blade_col = mix(base * np.array([0.78, 0.90, 0.82]), base * np.array([1.08, 1.16, 0.92]) + np.array([0.02, 0.04, 0.01]),
                    grad)
    out = mix(base * 0.85, blade_col, cov) * mix(0.68, 1.0, grad + 0.25 * cov)
    return np.maximum(out, 0.02)


def clip_grass_variation(hp, flags):
    """apply_material_variation_clip (clipmap.wgsl:247-275) for MAT_GRASS."""
    def var(size, strength):
        q = hp / size
        f = q - np.floor(q)
        flags["near"] |= np.any(np.minimum(f, 1.0 - f) < 1e-4, -1)
        c = np.floor(q).astype(np.int64)
        return (hash31_cells(c[:, 0], c[:, 1], c[:, 2]) - 0.5) * strength
    v = var(3.0, 1.0) + var(0.75, 0.35)
    c = (np.array([0.18, 0.75, 0.18])[None, :] + np.stack([0.10 * v, 0.18 * v, 0.06 * v], -1)) * (1.0 + 0.06 * v)[:, None]
    return clamp(c, 0.0, 2.0)


def shade_clip_hit_flat(ro, rd, t, time_s, flags):
    """shade_clip_hit (shading.wgsl:364-434) for a grass heightfield hit with
    normal +y on a flat clipmap (the AO-lite term is then 1)."""
    hp = ro[None, :] + rd * t[:, None]
    base = np.array([0.18, 0.75, 0.18])[None, :].repeat(len(t), 0)
    near = t <= 64.0
    base[near] = clip_grass_variation(hp[near], flags_sub := {"near": np.zeros(near.sum(), bool)})
    flags["near"][near] |= flags_sub["near"]
    base = grass_top_albedo(base, hp, t, flags)
    n = np.array([0.0, 1.0, 0.0])
    vis = cloud_sun_transmittance_fast(hp + n[None, :] * 0.075, time_s)
    diff = max(float(n @ SUN_DIR), 0.0)
    sky_up = sky_bg(np.array([[0.0, 1.0, 0.0]]))[0]
    ambient = mix(FOG_COLOR_GROUND, sky_up, 1.0) * 0.10
    direct = (SUN_COLOR * SUN_INTENSITY)[None, :] * (diff * diff * vis)[:, None]
    v = -rd / np.linalg.norm(rd, axis=-1, keepdims=True)
    h = v + SUN_DIR[None, :]
    h = h / np.linalg.norm(h, axis=-1, keepdims=True)
    ndv, ndh = np.maximum(v @ n, 0.0), np.maximum(h @ n, 0.0)
    fres = 0.03 + 0.97 * (1.0 - clamp(ndv, 0.0, 1.0)) ** 5.0
    spec = (SUN_COLOR * SUN_INTENSITY)[None, :] * (ndh ** mix(8.0, 96.0, 1.0 - 0.88) * fres * vis)[:, None]
    spec = np.where(near[:, None], spec, 0.0)
    return base * (ambient[None, :] + direct) + 0.18 * spec
