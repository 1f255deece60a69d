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
