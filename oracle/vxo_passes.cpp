// vxo_passes.cpp — oracle restatement of ray/shading.wgsl, clipmap.wgsl,
// ray/godrays.wgsl, ray/composite.wgsl, taa.wgsl and the compute entry points
// of ray_main.wgsl (main_primary, main_godray, main_composite), with a C ABI
// for the tests, the smoke check and the CPU baseline leg of the benchmark.
//
// TEST INFRASTRUCTURE ONLY.  This is the checker of the HIP passes.  It is
// never linked into, loaded by, or called from the product library.
//
// How each part is pinned independently of this file:
//   traversal, queries        dense brute-force voxel walks (test_traversal.py,
//                             test_query_oracle.py)
//   composite, both TAAs, the godray history blend, the cloud-free sky
//                             float64 statements (test_oracle_composition.py)
//   sky cloud march, godray march with fog on (empty world and a slab roof),
//   dirt/stone/wood voxel shading, grass heightfield shading
//                             float64 statements (test_oracle_shading.py)
//   leaf helpers              closed forms (test_oracle_math.py)
// Parity unpinned: the wind field away from the ground (hash1 of a sin, so its
// value follows the sin implementation), the voxel-grass shading extras
// (wrapped diffuse, back-scatter, grass self-shadow), the leaf dapple and the
// displaced leaf cubes; the exact transcendental values.
#include <algorithm>
#include <atomic>
#include <thread>
#include <vector>

#include "vxo_trace.h"

namespace vxo {

// ===========================================================================
// ray/shading.wgsl
// ===========================================================================
// color_for_material, shading.wgsl:7-16
static inline v3 color_for_material(uint32_t m) {
  if (m == MAT_AIR) return V3s(0.0f);
  if (m == MAT_GRASS) return V3(0.18f, 0.75f, 0.18f);
  if (m == MAT_DIRT) return V3(0.45f, 0.30f, 0.15f);
  if (m == MAT_STONE) return V3(0.50f, 0.50f, 0.55f);
  if (m == MAT_WOOD) return V3(0.38f, 0.26f, 0.14f);
  if (m == MAT_LEAF) return V3(0.10f, 0.55f, 0.12f);
  if (m == MAT_LIGHT) return V3(1.0f, 0.95f, 0.75f);
  return V3(1.0f, 0.0f, 1.0f);
}
// hemi_ambient, shading.wgsl:18-22
static inline v3 hemi_ambient(v3 n, v3 sky_up) {
  float upw = clampf(n.y * 0.5f + 0.5f, 0.0f, 1.0f);
  return mix3(FOG_COLOR_GROUND, sky_up, upw);
}
// material_variation / apply_material_variation, shading.wgsl:24-50
static inline v3 apply_material_variation(v3 base, uint32_t mat, v3 hp) {
  v3 c = base;
  v3 cell = floor3(hp / V3s(0.05f));
  float v = ALBEDO_VAR_GAIN * ((hash31(cell) - 0.5f) * 2.0f);
  if (mat == MAT_GRASS) {
    c = c + V3(0.02f * v, 0.05f * v, 0.01f * v);
    c = c * (1.0f + 0.06f * v);
  } else if (mat == MAT_DIRT) {
    c = c + V3(0.04f * v, 0.02f * v, 0.01f * v);
    c = c * (1.0f + 0.08f * v);
  } else if (mat == MAT_STONE) {
    c = c * (1.0f + 0.10f * v);
  } else if (mat == MAT_WOOD) {
    c = c + V3(0.05f * v, 0.02f * v, 0.00f * v);
    c = c * (1.0f + 0.07f * v);
  } else if (mat == MAT_LEAF) {
    c = c + V3(0.00f * v, 0.03f * v, 0.00f * v);
    c = c * (1.0f + 0.04f * v);
  }
  return clamp3s(c, 0.0f, 1.5f);
}
// grass_wind_dir_xz, shading.wgsl:52-59
static inline v2 grass_wind_dir_xz(v3 p, float time_s) {
  v3 w3 = wind_field(p, time_s);
  v2 dir = V2(w3.x, w3.z) + V2(0.6f, 0.2f);
  if (dot2(dir, dir) < 1e-4f) dir = V2(0.7f, 0.25f);
  return normalize2(dir);
}
// grass_blade_slivers, shading.wgsl:61-88
static inline float grass_blade_slivers(v3 p_ws, float time_s, float freq, float detail) {
  v2 dir = grass_wind_dir_xz(p_ws, time_s);
  v2 perp = V2(-dir.y, dir.x);
  v2 pxz = V2(p_ws.x, p_ws.z);
  v2 p = V2(dot2(pxz, dir), dot2(pxz, perp)) * freq;
  v2 cell = floor2(p);
  v2 f = fract2(p);
  float h0 = hash31(V3(cell.x, cell.y, 11.3f));
  float h1 = hash31(V3(cell.x + 3.7f, cell.y + 7.1f, 5.9f));
  float angle = (h0 - 0.5f) * 0.35f;
  float ca = vx_cos(angle), sa = vx_sin(angle);
  v2 pr = V2(f.x * ca - f.y * sa, f.x * sa + f.y * ca);
  float center = 0.5f + (h1 - 0.5f) * 0.6f;
  float u = clampf(pr.x, 0.0f, 1.0f);
  float taper = mixf(0.065f, 0.012f, vx_pow(u, 1.8f));
  float width = taper * mixf(0.75f, 1.25f, h0);
  float aa = mixf(1.6f, 0.6f, detail) / freq;
  float dist = std::fabs(pr.y - center);
  float sliver = smoothstepf(width + aa, width - aa, dist);
  float length_mask = smoothstepf(0.0f, 0.15f, u) * smoothstepf(1.0f, 0.75f, u);
  return sliver * length_mask;
}
// grass_blade_field, shading.wgsl:90-98
static inline float grass_blade_field(v3 p, float time_s, float t) {
  float detail = 1.0f - smoothstepf(18.0f, 55.0f, t);
  float low = grass_blade_slivers(p, time_s, 9.0f, detail * 0.6f);
  float high = grass_blade_slivers(p + V3(1.7f, 0.0f, 2.3f), time_s, 18.0f, detail);
  float blend = mixf(0.15f, 1.0f, detail);
  return mixf(low, 0.5f * (low + high), blend);
}
// grass_top_albedo, shading.wgsl:100-122
static inline v3 grass_top_albedo(const Ctx& C, v3 base, v3 hp, float t) {
  float vs = vp_x(C);
  float tip = clampf(fractf(hp.y / mx(vs, 1e-6f)), 0.0f, 1.0f);
  float time_s = C.cam->voxel_params[1];
  float blade = grass_blade_field(hp, time_s, t);
  float patches = fbm(V2(hp.x, hp.z) * 0.08f);
  float clumps = fbm(V2(hp.x, hp.z) * 0.35f + V2(12.3f, 7.1f));
  float density = smoothstepf(0.25f, 0.75f, patches) * smoothstepf(0.20f, 0.80f, clumps);
  float coverage = blade * density;
  v3 cool_base = base * V3(0.78f, 0.90f, 0.82f);
  v3 warm_tip = base * V3(1.08f, 1.16f, 0.92f) + V3(0.02f, 0.04f, 0.01f);
  float grad = smoothstepf(0.0f, 1.0f, tip);
  v3 blade_col = mix3(cool_base, warm_tip, grad);
  v3 out = mix3(base * 0.85f, blade_col, coverage);
  float base_shadow = mixf(0.68f, 1.0f, grad + 0.25f * coverage);
  out = out * base_shadow;
  return mx3(out, V3s(0.02f));
}
// base_albedo, shading.wgsl:124-130
static inline v3 base_albedo(uint32_t mat, v3 hp, float t) {
  v3 base = color_for_material(mat);
  if (t <= FAR_SHADING_DIST) base = apply_material_variation(base, mat, hp);
  return base;
}
// macro_occ_at_ws, shading.wgsl:132-154
static inline float macro_occ_at_ws(const Ctx& C, v3 p, v3 root_bmin, float root_size, uint32_t macro_base) {
  if (macro_base == INVALID_U32) return 0.0f;
  v3 lp = p - root_bmin;
  if (lp.x < 0.0f || lp.y < 0.0f || lp.z < 0.0f || lp.x >= root_size || lp.y >= root_size || lp.z >= root_size) return 0.0f;
  float cell = root_size / (float)MACRO_DIM;
  uint32_t ux = f2u(std::floor(lp.x / cell)), uy = f2u(std::floor(lp.y / cell)), uz = f2u(std::floor(lp.z / cell));
  ux = ux < MACRO_DIM - 1u ? ux : MACRO_DIM - 1u;
  uy = uy < MACRO_DIM - 1u ? uy : MACRO_DIM - 1u;
  uz = uz < MACRO_DIM - 1u ? uz : MACRO_DIM - 1u;
  return macro_test(C, macro_base, macro_bit_index(ux, uy, uz)) ? 1.0f : 0.0f;
}
// voxel_ao_local, shading.wgsl:157-184
static inline float voxel_ao_local(const Ctx& C, v3 hp, v3 n, v3 root_bmin, float root_size, uint32_t macro_base) {
  float r = 0.75f * vp_x(C);
  v3 up_ref = std::fabs(n.y) > 0.9f ? V3(1.0f, 0.0f, 0.0f) : V3(0.0f, 1.0f, 0.0f);
  v3 t = normalize3(cross3(up_ref, n));
  v3 b = normalize3(cross3(n, t));
  float occ = 0.0f;
  occ += macro_occ_at_ws(C, hp + t * r, root_bmin, root_size, macro_base);
  occ += macro_occ_at_ws(C, hp - t * r, root_bmin, root_size, macro_base);
  occ += macro_occ_at_ws(C, hp + b * r, root_bmin, root_size, macro_base);
  occ += macro_occ_at_ws(C, hp - b * r, root_bmin, root_size, macro_base);
  float occ_n = occ * 0.25f;
  return clampf(1.0f - 0.70f * occ_n, 0.35f, 1.0f);
}
// fresnel_schlick / material_roughness / material_f0, shading.wgsl:187-209
static inline float fresnel_schlick(float ndv, float f0) { return f0 + (1.0f - f0) * vx_pow(1.0f - clampf(ndv, 0.0f, 1.0f), 5.0f); }
static inline float material_roughness(uint32_t m) {
  if (m == MAT_STONE) return 0.45f;
  if (m == MAT_WOOD) return 0.70f;
  if (m == MAT_LEAF) return 0.80f;
  if (m == MAT_GRASS) return 0.85f;
  if (m == MAT_DIRT) return 0.90f;
  if (m == MAT_LIGHT) return 0.25f;
  return 0.90f;
}
static inline float material_f0(uint32_t m) {
  if (m == MAT_STONE) return 0.04f;
  if (m == MAT_WOOD) return 0.03f;
  if (m == MAT_LEAF) return 0.05f;
  if (m == MAT_GRASS) return 0.04f;
  if (m == MAT_DIRT) return 0.02f;
  if (m == MAT_LIGHT) return 0.08f;
  return 0.02f;
}
// shade_hit, shading.wgsl:257-362
static inline v3 shade_hit(const Ctx& C, v3 ro, v3 rd, const HitGeom& hg, v3 sky_up, float sun_shadow) {
  v3 hp = ro + hg.t * rd;
  v3 base = base_albedo(hg.mat, hp, hg.t);
  if (hg.mat == MAT_GRASS) {
    base = grass_top_albedo(C, base, hp, hg.t);
    float ndl = dot3(hg.n, SUN_DIR);
    const float wrap = 0.35f;
    float diff_wrap = clampf((ndl + wrap) / (1.0f + wrap), 0.0f, 1.0f);
    float back = vx_pow(clampf(dot3(-SUN_DIR, hg.n), 0.0f, 1.0f), 1.6f);
    float view_graze = vx_pow(1.0f - mx(dot3(normalize3(-rd), hg.n), 0.0f), 2.0f);
    base = base + (0.22f * back + 0.10f * view_graze * back) * V3(0.18f, 0.34f, 0.10f);
    base = base * (0.90f + 0.10f * diff_wrap);
  }
  float vs = vp_x(C);
  v3 hp_shadow = hp + hg.n * (0.75f * vs);
  float Tc = cloud_sun_transmittance_fast(C, hp_shadow, SUN_DIR);
  float sun_vis = Tc * clampf(sun_shadow, 0.0f, 1.0f);
  float diff = mx(dot3(hg.n, SUN_DIR), 0.0f);
  float ao = 1.0f;
  if (hg.hit != 0u && hg.t <= VOXEL_AO_MAX_DIST) ao = voxel_ao_local(C, hp, hg.n, hg.root_bmin, hg.root_size, hg.macro_base);
  v3 amb_col = hemi_ambient(hg.n, sky_up);
  float amb_strength = hg.mat == MAT_LEAF ? 0.13f : 0.10f;
  float ao_term = vx_pow(clampf(ao, 0.0f, 1.0f), 1.0f);
  v3 ambient = amb_col * amb_strength * ao_term;
  if (hg.mat == MAT_STONE) ambient = ambient * V3(0.92f, 0.95f, 1.05f);
  float dapple = 1.0f;
  if (hg.mat == MAT_LEAF && hg.t <= FAR_SHADING_DIST) {
    float time_s = C.cam->voxel_params[1];
    v2 hxz = V2(hp.x, hp.z);
    float d0 = vx_sin(dot2(hxz, V2(3.0f, 2.2f)) + time_s * 3.5f);
    float d1 = vx_sin(dot2(hxz, V2(6.5f, 4.1f)) - time_s * 6.0f);
    dapple = 0.90f + 0.10f * (0.6f * d0 + 0.4f * d1);
  }
  v3 spec_col = V3s(0.0f);
  if (hg.t <= FAR_SHADING_DIST) {
    v3 v = normalize3(-rd);
    v3 h = normalize3(v + SUN_DIR);
    float ndv = mx(dot3(hg.n, v), 0.0f);
    float ndh = mx(dot3(hg.n, h), 0.0f);
    float rough = material_roughness(hg.mat);
    float shininess = mixf(8.0f, 96.0f, 1.0f - rough);
    float spec = vx_pow(ndh, shininess);
    float f0 = material_f0(hg.mat);
    float fres = fresnel_schlick(ndv, f0);
    spec_col = SUN_COLOR * SUN_INTENSITY * spec * fres * sun_vis;
  }
  float grass_ts = 1.0f;
  if (hg.mat == MAT_GRASS) {
    float time_s = C.cam->voxel_params[1], strength = C.cam->voxel_params[2];
    GrassCell cell = grass_cell_from_world(hp, rd, hg.root_bmin, vs, (int32_t)C.cam->chunk_size);
    uint32_t lod = grass_lod_from_t(hg.t);
    grass_ts = grass_self_shadow(C, hp, cell.bmin_m, cell.id_vox, time_s, strength, lod);
  }
  v3 direct = SUN_COLOR * SUN_INTENSITY * (0.35f * diff + 0.65f * diff * diff) * sun_vis * dapple * grass_ts;
  v3 emissive = hg.mat == MAT_LIGHT ? 18.0f * V3(1.0f, 0.95f, 0.75f) : V3s(0.0f);
  return base * (ambient + direct) + 0.20f * spec_col + emissive;
}
struct ShadeOut { v3 base_hdr, local_hdr; float local_w; };
// shade_hit_split, shading.wgsl:224-255
static inline ShadeOut shade_hit_split(const Ctx& C, v3 ro, v3 rd, const HitGeom& hg, v3 sky_up, float sun_shadow) {
  v3 base_hdr = shade_hit(C, ro, rd, hg, sky_up, sun_shadow);
  v3 local_hdr = V3s(0.0f);
  float local_w = 0.0f;
  if (hg.hit != 0u && hg.mat != MAT_LIGHT && hg.t <= LOCAL_LIGHT_MAX_DIST) {
    v3 hp = ro + hg.t * rd;
    v3 albedo = base_albedo(hg.mat, hp, hg.t);
    local_hdr = gather_voxel_lights(C, hp, hg.n, hg.root_bmin);
    local_hdr = local_hdr * albedo;
    local_w = 1.0f;
  }
  return ShadeOut{base_hdr, local_hdr, local_w};
}

// ===========================================================================
// clipmap.wgsl
// ===========================================================================
static inline int32_t imod(int32_t a, int32_t m) {
  int32_t r = a % m;
  if (r < 0) r = r + m;
  return r;
}
static inline int32_t clip_res_i(const Ctx& C) { return C.clip->res > 1u ? (int32_t)C.clip->res : 1; }
// clip_st, clipmap.wgsl:317-327
static inline v2 clip_st(const Ctx& C, v2 xz, uint32_t level) {
  const float* p = C.clip->level[level];
  v2 origin = V2(p[0], p[1]);
  float cell = mx(p[2], 1e-6f);
  v2 uv = (xz - origin) / V2(cell, cell);
  return uv - V2(0.5f, 0.5f);
}
// clip_height_texel, clipmap.wgsl:55-67
static inline float clip_height_texel(const Ctx& C, uint32_t level, int32_t ix, int32_t iz) {
  int32_t r = clip_res_i(C);
  if (ix < 0 || iz < 0 || ix >= r || iz >= r) return -BIG_F32;
  int32_t sx = imod(ix + (int32_t)C.clip->offset[level][0], r), sz = imod(iz + (int32_t)C.clip->offset[level][1], r);
  return ld_clip(C, level, sx, sz);
}
struct HSample { float h; bool ok; };
// clip_height_at_level_ok, clipmap.wgsl:34-52
static inline HSample clip_height_at_level_ok(const Ctx& C, v2 xz, uint32_t level) {
  int32_t r = clip_res_i(C);
  v2 st = clip_st(C, xz, level);
  int32_t ix = f2i(std::floor(st.x)), iz = f2i(std::floor(st.y));
  if (ix < 0 || iz < 0 || ix >= r || iz >= r) return HSample{0.0f, false};
  int32_t sx = imod(ix + (int32_t)C.clip->offset[level][0], r), sz = imod(iz + (int32_t)C.clip->offset[level][1], r);
  return HSample{ld_clip(C, level, sx, sz), true};
}
// clip_height_at_level (bilinear), clipmap.wgsl:71-108
static inline float clip_height_at_level(const Ctx& C, v2 xz, uint32_t level) {
  int32_t r = clip_res_i(C);
  const float* p = C.clip->level[level];
  v2 origin = V2(p[0], p[1]);
  float cell = mx(p[2], 1e-6f);
  v2 uv = (xz - origin) / V2(cell, cell);
  v2 st = uv - V2(0.5f, 0.5f);
  int32_t ix0 = f2i(std::floor(st.x)), iz0 = f2i(std::floor(st.y));
  int32_t ix1 = ix0 + 1, iz1 = iz0 + 1;
  if (ix0 < 0 || iz0 < 0 || ix1 >= r || iz1 >= r) return -BIG_F32;
  float fx = fractf(st.x), fz = fractf(st.y);
  float h00 = clip_height_texel(C, level, ix0, iz0);
  float h10 = clip_height_texel(C, level, ix1, iz0);
  float h01 = clip_height_texel(C, level, ix0, iz1);
  float h11 = clip_height_texel(C, level, ix1, iz1);
  if (h00 <= -0.5f * BIG_F32 || h10 <= -0.5f * BIG_F32 || h01 <= -0.5f * BIG_F32 || h11 <= -0.5f * BIG_F32) return -BIG_F32;
  float hx0 = mixf(h00, h10, fx), hx1 = mixf(h01, h11, fx);
  return mixf(hx0, hx1, fz);
}
// clip_height_at_level_nearest, clipmap.wgsl:110-130
static inline float clip_height_at_level_nearest(const Ctx& C, v2 xz, uint32_t level) {
  int32_t r = clip_res_i(C);
  const float* p = C.clip->level[level];
  v2 origin = V2(p[0], p[1]);
  float cell = mx(p[2], 1e-6f);
  v2 uv = (xz - origin) / V2(cell, cell);
  v2 st = uv - V2(0.5f, 0.5f);
  int32_t ix = f2i(std::floor(st.x + 0.5f)), iz = f2i(std::floor(st.y + 0.5f));
  int32_t sx = imod(ix + (int32_t)C.clip->offset[level][0], r), sz = imod(iz + (int32_t)C.clip->offset[level][1], r);
  return ld_clip(C, level, sx, sz);
}
// clip_normal_at_level_2tap, clipmap.wgsl:132-147
static inline v3 clip_normal_at_level_2tap(const Ctx& C, v2 xz, uint32_t level) {
  float cell = C.clip->level[level][2];
  HSample h0 = clip_height_at_level_ok(C, xz, level);
  HSample hx = clip_height_at_level_ok(C, xz + V2(cell, 0.0f), level);
  HSample hz = clip_height_at_level_ok(C, xz + V2(0.0f, cell), level);
  if (!h0.ok || !hx.ok || !hz.ok) return V3(0.0f, 1.0f, 0.0f);
  float dhx = (hx.h - h0.h) / mx(cell, 1e-4f);
  float dhz = (hz.h - h0.h) / mx(cell, 1e-4f);
  return normalize3(V3(-dhx, 1.0f, -dhz));
}
// clip_level_contains / clip_best_level / clip_ensure_contains, clipmap.wgsl:277-315
static inline bool clip_level_contains(const Ctx& C, v2 xz, uint32_t level, int32_t guard) {
  int32_t r = clip_res_i(C);
  v2 st = clip_st(C, xz, level);
  int32_t ix0 = f2i(std::floor(st.x)), iz0 = f2i(std::floor(st.y));
  int32_t ix1 = ix0 + 1, iz1 = iz0 + 1;
  return (ix0 >= guard) && (iz0 >= guard) && (ix1 < (r - guard)) && (iz1 < (r - guard));
}
static inline uint32_t clip_nlevels(const Ctx& C) { return C.clip->levels < CLIP_LEVELS_MAX ? C.clip->levels : CLIP_LEVELS_MAX; }
static inline uint32_t clip_best_level(const Ctx& C, v2 xz, int32_t guard) {
  uint32_t n = clip_nlevels(C);
  for (uint32_t i = 0; i < n; ++i)
    if (clip_level_contains(C, xz, i, guard)) return i;
  return (n > 1u ? n : 1u) - 1u;
}
static inline uint32_t clip_ensure_contains(const Ctx& C, v2 xz, uint32_t lvl_in, int32_t guard) {
  uint32_t n = clip_nlevels(C);
  if (n == 0u) return 0u;
  uint32_t lvl = lvl_in < n - 1u ? lvl_in : n - 1u;
  for (;;) {
    if (clip_level_contains(C, xz, lvl, guard) || lvl >= (n - 1u)) break;
    lvl = lvl + 1u;
  }
  return lvl;
}
struct ClipHit { bool hit; float t; v3 n; uint32_t mat; };
static inline ClipHit clip_miss() { return ClipHit{false, BIG_F32, V3s(0.0f), MAT_AIR}; }
// clip_trace_heightfield, clipmap.wgsl:156-243
static inline ClipHit clip_trace_heightfield(const Ctx& C, v3 ro, v3 rd, float t_min, float t_max) {
  if (C.clip->levels == 0u) return clip_miss();
  if (rd.y >= -1e-4f) return clip_miss();
  const int32_t guard = 2;
  float t = mx(t_min, 0.0f);
  v3 p = ro + rd * t;
  uint32_t lvl = clip_best_level(C, V2(p.x, p.z), guard);
  float h0 = clip_height_at_level_nearest(C, V2(p.x, p.z), lvl);
  if (h0 <= -0.5f * BIG_F32) return clip_miss();
  float s_prev = p.y - h0;
  float t_prev = t;
  for (uint32_t i = 0; i < HF_MAX_STEPS; ++i) {
    if (t > t_max) break;
    p = ro + rd * t;
    lvl = clip_ensure_contains(C, V2(p.x, p.z), lvl, guard);
    float h = clip_height_at_level_nearest(C, V2(p.x, p.z), lvl);
    if (h <= -0.5f * BIG_F32) return clip_miss();
    float s = p.y - h;
    if (s <= 0.0f && s_prev > 0.0f) {
      float a = t_prev, b = t;
      for (uint32_t k = 0; k < HF_BISECT; ++k) {
        float m = 0.5f * (a + b);
        v3 pm = ro + rd * m;
        float hm = clip_height_at_level(C, V2(pm.x, pm.z), lvl);
        if (hm <= -0.5f * BIG_F32) return clip_miss();
        float sm = pm.y - hm;
        if (sm > 0.0f) a = m;
        else b = m;
      }
      float th = 0.5f * (a + b);
      v3 ph = ro + rd * th;
      v3 n = clip_normal_at_level_2tap(C, V2(ph.x, ph.z), lvl);
      return ClipHit{true, th, n, MAT_GRASS};
    }
    s_prev = s;
    t_prev = t;
    float vy = mx(-rd.y, 0.12f);
    float vh = mx(length2(V2(rd.x, rd.z)), 1e-4f);
    float dt_y = std::fabs(s) / vy;
    float cell = C.clip->level[lvl][2];
    float dt_xz = (6.0f * cell) / vh;
    float dt = mn(dt_y, dt_xz);
    dt = clampf(dt, 0.6f, HF_DT_MAX);
    t = t + dt;
  }
  return clip_miss();
}
// apply_material_variation_clip, clipmap.wgsl:247-275
static inline float material_variation_clip(v3 p, float cs, float strength) {
  v3 cell = floor3(p / V3s(cs));
  return (hash31(cell) - 0.5f) * strength;
}
static inline v3 apply_material_variation_clip(v3 base, uint32_t mat, v3 hp) {
  v3 c = base;
  if (mat == MAT_GRASS) {
    float v0 = material_variation_clip(hp, 3.0f, 1.0f);
    float v1 = material_variation_clip(hp, 0.75f, 0.35f);
    float v = v0 + v1;
    c = c + V3(0.10f * v, 0.18f * v, 0.06f * v);
    c = c * (1.0f + 0.06f * v);
    c = clamp3s(c, 0.0f, 2.0f);
  } else if (mat == MAT_DIRT) {
    float v = material_variation_clip(hp, 1.5f, 0.8f);
    c = c + V3(0.05f * v, 0.03f * v, 0.01f * v);
    c = c * (1.0f + 0.08f * v);
    c = clamp3s(c, 0.0f, 2.0f);
  } else if (mat == MAT_STONE) {
    float v = material_variation_clip(hp, 2.0f, 0.9f);
    c = c * (1.0f + 0.10f * v);
    c = clamp3s(c, 0.0f, 2.0f);
  }
  return c;
}
// shade_clip_hit, shading.wgsl:364-434
static inline v3 shade_clip_hit(const Ctx& C, v3 ro, v3 rd, const ClipHit& ch, v3 sky_up) {
  v3 hp = ro + ch.t * rd;
  v3 base = color_for_material(ch.mat);
  if (ch.t <= FAR_SHADING_DIST) base = apply_material_variation_clip(base, ch.mat, hp);
  if (ch.mat == MAT_GRASS) base = grass_top_albedo(C, base, hp, ch.t);
  float vs = vp_x(C);
  v3 hp_shadow = hp + ch.n * (0.75f * vs);
  float vis = cloud_sun_transmittance_fast(C, hp_shadow, SUN_DIR);
  float diff = mx(dot3(ch.n, SUN_DIR), 0.0f);
  float ao = 1.0f;
  if (ch.mat == MAT_GRASS && ch.t <= FAR_SHADING_DIST) {
    uint32_t lvl = clip_best_level(C, V2(hp.x, hp.z), 2);
    float cell = C.clip->level[lvl][2];
    v2 hxz = V2(hp.x, hp.z);
    float h0 = clip_height_at_level(C, hxz, lvl);
    float hx1 = clip_height_at_level(C, hxz + V2(cell, 0.0f), lvl);
    float hx0 = clip_height_at_level(C, hxz + V2(-cell, 0.0f), lvl);
    float hz1 = clip_height_at_level(C, hxz + V2(0.0f, cell), lvl);
    float hz0 = clip_height_at_level(C, hxz + V2(0.0f, -cell), lvl);
    float occ = mx(0.0f, hx1 - h0) + mx(0.0f, hx0 - h0) + mx(0.0f, hz1 - h0) + mx(0.0f, hz0 - h0);
    ao = clampf(1.0f - 0.65f * occ / mx(cell, 1e-3f), 0.45f, 1.0f);
  }
  v3 amb_col = hemi_ambient(ch.n, sky_up);
  const float amb_strength = 0.10f;
  v3 ambient = amb_col * amb_strength * ao;
  v3 direct = SUN_COLOR * SUN_INTENSITY * (diff * diff) * vis;
  v3 spec_col = V3s(0.0f);
  if (ch.t <= FAR_SHADING_DIST) {
    v3 vdir = normalize3(-rd);
    v3 hdir = normalize3(vdir + SUN_DIR);
    float ndv = mx(dot3(ch.n, vdir), 0.0f);
    float ndh = mx(dot3(ch.n, hdir), 0.0f);
    float rough = 0.85f;
    if (ch.mat == MAT_STONE) rough = 0.50f;
    if (ch.mat == MAT_DIRT) rough = 0.90f;
    if (ch.mat == MAT_GRASS) rough = 0.88f;
    float shininess = mixf(8.0f, 96.0f, 1.0f - rough);
    float spec = vx_pow(ndh, shininess);
    float f0 = 0.03f;
    if (ch.mat == MAT_STONE) f0 = 0.04f;
    float fres = f0 + (1.0f - f0) * vx_pow(1.0f - clampf(ndv, 0.0f, 1.0f), 5.0f);
    spec_col = SUN_COLOR * SUN_INTENSITY * spec * fres * vis;
  }
  return base * (ambient + direct) + 0.18f * spec_col;
}

// ===========================================================================
// ray_main.wgsl helpers
// ===========================================================================
static inline uint32_t pack_i16x2(int32_t a, int32_t b) { return ((uint32_t)a & 0xFFFFu) | (((uint32_t)b & 0xFFFFu) << 16u); }
static inline int32_t unpack_i16(uint32_t v) {
  int32_t s = (int32_t)(v & 0xFFFFu);
  return (v & 0x8000u) != 0u ? s - 0x10000 : s;
}
// profile_inc / profile_add, ray_main.wgsl:64-74 (counters of ray_main.wgsl:32-47)
enum { PROFILE_VOXEL = 0, PROFILE_GRASS = 1, PROFILE_HDR = 2, PROFILE_FOG = 3, PROFILE_SHADOW = 4 };
static inline void profile_add(const Ctx& C, uint32_t idx, uint32_t value) {
  if ((C.cam->profile_flags & 1u) != 0u) C.cnt->prof[idx] += value;
}
// fog_profile_weight, ray_main.wgsl:90-94 (ENABLE_FOG is true)
static inline uint32_t fog_profile_weight(float t_scene) {
  float fog_norm = clampf(t_scene / FOG_MAX_DIST, 0.0f, 1.0f);
  uint32_t w = f2u(fog_norm * 256.0f);
  return w > 1u ? w : 1u;
}
struct HfShadeOut { v3 col; float t_scene; };
// shade_heightfield, ray_main.wgsl:100-114
static inline HfShadeOut shade_heightfield(const Ctx& C, v3 ro, v3 rd, const ClipHit& hf, v3 sky_up) {
  profile_add(C, PROFILE_HDR, 1u);
  v3 surface = shade_clip_hit(C, ro, rd, hf, sky_up);
  float t_scene = mn(hf.t, FOG_MAX_DIST);
  v3 sbg = sky_bg(rd);
  v3 col = apply_fog(C, surface, ro, rd, t_scene, sbg);
  profile_add(C, PROFILE_FOG, fog_profile_weight(t_scene));
  return HfShadeOut{col, t_scene};
}
// trace_primary_voxels, ray_main.wgsl:116-161
static inline VoxTraceResult trace_primary_voxels(const Ctx& C, const TileList& L, v3 ro, v3 rd, float t_start, float t_end,
                                                  bool has_tile, bool hist_valid, uint32_t hak, i3 hac, uint32_t tcc) {
  VoxTraceResult vt{false, miss_hitgeom(), 0.0f, false, INVALID_U32, i3{0, 0, 0}};
  if (has_tile) {
    if (hist_valid) {
      profile_add(C, PROFILE_VOXEL, tcc > 1u ? tcc : 1u);
      VoxTraceResult vh = trace_scene_voxels_candidates(C, L, ro, rd, t_start, t_end, hak != INVALID_U32, hac, hak, tcc);
      if (vh.best.hit != 0u) return vh;
    }
    profile_add(C, PROFILE_VOXEL, tcc > 1u ? tcc : 1u);
    vt = trace_scene_voxels_candidates(C, L, ro, rd, 0.0f, FOG_MAX_DIST, false, i3{0, 0, 0}, INVALID_U32, tcc);
  }
  return vt;
}

// Nearest/clamp-to-edge sampling (render/state/mod.rs:216-225): texel = clamp(floor(uv*dim)).
static inline int32_t nearest_texel(float u, uint32_t dim) { return clampi(f2i(std::floor(u * (float)dim)), 0, (int32_t)dim - 1); }

// Per-view textures of one frame (row-major, RGBA32F = 4 floats).
struct PrimaryIO {
  uint32_t w, h;
  const float* primary_hist_in;  // RGBA
  const float* shadow_hist_in;   // R
  float* color;                  // RGBA
  float* depth;                  // R
  float* local;                  // RGBA
  float* primary_hist_out;       // RGBA
  float* shadow_hist_out;        // R
  vx_hit* hits;                  // optional
};

static inline void store4(float* img, uint32_t w, int32_t x, int32_t y, float a, float b, float c, float d) {
  float* p = img + ((size_t)y * w + (size_t)x) * 4;
  p[0] = a;
  p[1] = b;
  p[2] = c;
  p[3] = d;
}

// main_primary, ray_main.wgsl:163-517, for one 8x8 workgroup.
static void main_primary_tile(const Ctx& C, const PrimaryIO& io, uint32_t wgx, uint32_t wgy) {
  const vx_camera* cam = C.cam;
  TileList L;
  L.count = 0;
  v3 sky_up = sky_bg(V3(0.0f, 1.0f, 0.0f));
  if (cam->chunk_count != 0u) {
    v2 tile_base = V2((float)(wgx * TILE_SIZE), (float)(wgy * TILE_SIZE));
    v3 ro_tile = cam_pos(C);
    v2 px = tile_base + V2(4.5f, 4.5f);
    tile_append_candidates_for_ray(C, L, ro_tile, ray_dir_from_pixel(C, px), 0.0f, FOG_MAX_DIST);
    px = tile_base + V2(1.5f, 1.5f);
    tile_append_candidates_for_ray(C, L, ro_tile, ray_dir_from_pixel(C, px), 0.0f, FOG_MAX_DIST);
  }
  uint32_t raw_count = L.count < MAX_TILE_CHUNKS ? L.count : MAX_TILE_CHUNKS;
  tile_sort_candidates_by_enter(L, raw_count);
  L.cached = raw_count < PRIMARY_MAX_TILE_CHUNKS ? raw_count : PRIMARY_MAX_TILE_CHUNKS;

  for (uint32_t ly = 0; ly < TILE_SIZE; ++ly)
    for (uint32_t lx = 0; lx < TILE_SIZE; ++lx) {
      uint32_t gx = wgx * TILE_SIZE + lx, gy = wgy * TILE_SIZE + ly;
      if (gx >= io.w || gy >= io.h) continue;
      v2 res = V2((float)io.w, (float)io.h);
      v2 px = V2((float)gx + 0.5f, (float)gy + 0.5f);
      uint32_t frame = cam->frame_index;
      uint32_t seed = (gx * 1973u) ^ (gy * 9277u) ^ (frame * 26699u);
      v3 ro = cam_pos(C);
      v3 rd = ray_dir_from_pixel(C, px);
      int32_t ipx = (int32_t)gx, ipy = (int32_t)gy;
      size_t pix = (size_t)gy * io.w + gx;
      v3 local_out = V3s(0.0f);
      float local_w = 0.0f, t_store = 0.0f, shadow_out = 1.0f;
      vx_hit hr;
      hr.slot = INVALID_U32;
      hr.node = INVALID_U32;
      hr.voxel[0] = hr.voxel[1] = hr.voxel[2] = 0;
      hr.mat = MAT_AIR;
      hr.t = BIG_F32;
      hr.kind = 0;

      auto finish_hf_or_sky = [&](ClipHit hf) {
        if (hf.hit) {
          if (hf.mat == MAT_GRASS) {
            float vs = vp_x(C);
            v3 hp = ro + rd * hf.t;
            GrassCell cell = grass_cell_from_world_global(hp, rd, vs);
            v3 gb = V3(cell.bmin_m.x, hp.y - vs, cell.bmin_m.z);
            GrassHit gh = try_grass_slab_hit(C, ro, rd, 0.0f, hf.t, gb, cell.id_vox, vs, cam->voxel_params[1], cam->voxel_params[2]);
            if (gh.hit) {
              hf.t = gh.t;
              hf.n = gh.n;
              profile_add(C, PROFILE_GRASS, 1u);
            }
          }
          HfShadeOut sh = shade_heightfield(C, ro, rd, hf, sky_up);
          t_store = sh.t_scene;
          store4(io.color, io.w, ipx, ipy, sh.col.x, sh.col.y, sh.col.z, 1.0f);
          io.depth[pix] = sh.t_scene;
          hr.kind = 2;
          hr.mat = hf.mat;
          hr.t = hf.t;
        } else {
          v3 sky = sky_color(C, rd);
          store4(io.color, io.w, ipx, ipy, sky.x, sky.y, sky.z, 1.0f);
          io.depth[pix] = FOG_MAX_DIST;
        }
        store4(io.local, io.w, ipx, ipy, local_out.x, local_out.y, local_out.z, local_w);
        store4(io.primary_hist_out, io.w, ipx, ipy, t_store, 0.0f, 0.0f, 0.0f);
        io.shadow_hist_out[pix] = shadow_out;
        C.cnt->tex_wr += 16 + 4 + 16 + 16 + 4;
        if (io.hits) io.hits[pix] = hr;
      };

      if (cam->chunk_count == 0u) {
        finish_hf_or_sky(clip_trace_heightfield(C, ro, rd, 0.0f, FOG_MAX_DIST));
        continue;
      }
      float t_hist = 0.0f;
      bool hist_valid = false;
      uint32_t hak = INVALID_U32;
      i3 hac{0, 0, 0};
      uint32_t tcc = L.cached;
      bool has_tile = tcc != 0u;
      v2 uv = px / res;
      if (has_tile) {
        const float* hg = io.primary_hist_in + pix * 4;
        C.cnt->tex_rd += 16;
        float t_hist_guess = hg[0];
        if (t_hist_guess > 1e-3f) {
          v3 p_ws = ro + rd * t_hist_guess;
          v2 uv_prev = prev_uv_from_world(C, p_ws);
          if (in_unit_square(uv_prev)) {
            int32_t sx = nearest_texel(uv_prev.x, io.w), sy = nearest_texel(uv_prev.y, io.h);
            const float* hp4 = io.primary_hist_in + ((size_t)sy * io.w + sx) * 4;
            C.cnt->tex_rd += 16;
            float t_prev = hp4[0];
            float rel = std::fabs(t_prev - t_hist_guess) / mx(t_hist_guess, 1e-3f);
            float depth_ok = 1.0f - smoothstepf(PRIMARY_HIT_DEPTH_REL0, PRIMARY_HIT_DEPTH_REL1, rel);
            float vel_px = length2((uv_prev - uv) * res);
            float motion_ok = 1.0f - smoothstepf(PRIMARY_HIT_MOTION_PX0, PRIMARY_HIT_MOTION_PX1, vel_px);
            if (t_prev > 1e-3f && depth_ok > 0.5f && motion_ok > 0.5f) {
              t_hist = t_prev;
              hist_valid = true;
              uint32_t packed_z = bits(hp4[3]);
              if ((packed_z & 0x80000000u) != 0u) {
                hak = bits(hp4[1]);
                uint32_t pxy = bits(hp4[2]);
                hac = i3{unpack_i16(pxy), unpack_i16(pxy >> 16u), unpack_i16(packed_z)};
              }
            }
          }
        }
      }
      float t_start = mx(t_hist - PRIMARY_HIT_MARGIN, 0.0f);
      float t_end = mn(t_hist + PRIMARY_HIT_WINDOW, FOG_MAX_DIST);
      VoxTraceResult vt = trace_primary_voxels(C, L, ro, rd, t_start, t_end, has_tile, hist_valid, hak, hac, tcc);
      if (!vt.in_grid || vt.best.hit == 0u) {
        finish_hf_or_sky(clip_trace_heightfield(C, ro, rd, 0.0f, FOG_MAX_DIST));
        continue;
      }
      // Voxel hit.
      HitGeom best = vt.best;
      if (best.mat == MAT_GRASS) {
        float vs = vp_x(C);
        int32_t csv = (int32_t)cam->chunk_size;
        i3 org{f2i(std::floor(best.root_bmin.x / vs + 0.5f)), f2i(std::floor(best.root_bmin.y / vs + 0.5f)),
               f2i(std::floor(best.root_bmin.z / vs + 0.5f))};
        v3 hp = ro + rd * best.t;
        GrassCell cell = pick_grass_cell_in_chunk_with_bias(hp, rd, best.root_bmin, org, vs, csv);
        GrassHit gh = try_grass_slab_hit(C, ro, rd, 0.0f, best.t, cell.bmin_m, cell.id_vox, vs, cam->voxel_params[1], cam->voxel_params[2]);
        if (gh.hit) {
          best.t = gh.t;
          best.n = gh.n;
          profile_add(C, PROFILE_GRASS, 1u);
        }
      }
      v3 hp = ro + best.t * rd;
      v3 hp_shadow = hp + best.n * (0.75f * vp_x(C));
      bool shadow_do = (seed & SHADOW_SUBSAMPLE_MASK) == 0u;
      if (shadow_do) {
        profile_add(C, PROFILE_SHADOW, 2u);
        float shadow_hist = io.shadow_hist_in[pix];
        C.cnt->tex_rd += 4;
        v2 uv_prev = prev_uv_from_world(C, hp);
        if (in_unit_square(uv_prev)) {
          int32_t ppx = clampi(f2i(uv_prev.x * (float)io.w), 0, (int32_t)io.w - 1);
          int32_t ppy = clampi(f2i(uv_prev.y * (float)io.h), 0, (int32_t)io.h - 1);
          shadow_hist = io.shadow_hist_in[(size_t)ppy * io.w + ppx];
          C.cnt->tex_rd += 4;
        }
        shadow_hist = clampf(shadow_hist, 0.0f, 1.0f);
        float shadow_cur = sun_transmittance_geom_only(C, hp_shadow, SUN_DIR);
        shadow_out = mixf(shadow_hist, shadow_cur, SHADOW_TAA_ALPHA);
      } else {
        profile_add(C, PROFILE_SHADOW, 1u);
        shadow_out = clampf(io.shadow_hist_in[pix], 0.0f, 1.0f);
        C.cnt->tex_rd += 4;
      }
      ShadeOut sh = shade_hit_split(C, ro, rd, best, sky_up, shadow_out);
      profile_add(C, PROFILE_HDR, 2u);
      float t_scene = mn(best.t, FOG_MAX_DIST);
      v3 sbg = sky_bg(rd);
      v3 col_base = apply_fog(C, sh.base_hdr, ro, rd, t_scene, sbg);
      profile_add(C, PROFILE_FOG, fog_profile_weight(t_scene));
      local_out = sh.local_hdr;
      local_w = sh.local_w;
      t_store = t_scene;
      store4(io.color, io.w, ipx, ipy, col_base.x, col_base.y, col_base.z, 1.0f);
      io.depth[pix] = t_scene;
      store4(io.local, io.w, ipx, ipy, local_out.x, local_out.y, local_out.z, local_w);
      uint32_t pxy = 0, pz = 0, ak = 0;
      if (vt.anchor_valid) {
        pxy = pack_i16x2(vt.anchor_chunk.x, vt.anchor_chunk.y);
        pz = pack_i16x2(vt.anchor_chunk.z, 0) | 0x80000000u;
        ak = vt.anchor_key;
      }
      store4(io.primary_hist_out, io.w, ipx, ipy, t_store, from_bits(ak), from_bits(pxy), from_bits(pz));
      io.shadow_hist_out[pix] = shadow_out;
      C.cnt->tex_wr += 16 + 4 + 16 + 16 + 4;
      if (io.hits) {
        hr.kind = 1;
        hr.slot = best.slot;
        hr.node = best.node;
        hr.voxel[0] = best.vox[0];
        hr.voxel[1] = best.vox[1];
        hr.voxel[2] = best.vox[2];
        hr.mat = best.mat;
        hr.t = best.t;
        io.hits[pix] = hr;
      }
    }
}

// ===========================================================================
// ray/godrays.wgsl
// ===========================================================================
static inline float approx_pow_0p75(float x) {
  float y = clampf(x, 0.0f, 1.0f);
  return mixf(y, std::sqrt(y), 0.5f);
}
static inline float approx_pow_0p55(float x) {
  float y = clampf(x, 0.0f, 1.0f);
  float s = std::sqrt(y);
  return mixf(s, y, 0.20f);
}
static inline float exp_neg_fast(float x_in) {
  float x = mx(x_in, 0.0f);
  float x2 = x * x;
  return 1.0f / (1.0f + x + 0.48f * x2);
}
// godray_steps_for_tend, godrays.wgsl:30-43
static inline uint32_t godray_steps_for_tend(float t_end) {
  float raw_f = std::ceil(t_end * GODRAY_STEPS_PER_METER);
  uint32_t s = f2u(clampf(raw_f, (float)GODRAY_MIN_STEPS, (float)GODRAY_STEPS_FAST));
  s = ((s + (GODRAY_STEP_Q - 1u)) / GODRAY_STEP_Q) * GODRAY_STEP_Q;
  return s < GODRAY_STEPS_FAST ? s : GODRAY_STEPS_FAST;
}
// godray_integrate_1, godrays.wgsl:45-216
static v3 godray_integrate_1(const Ctx& C, v3 ro, v3 rd, float t_end, float j, float j_phase) {
  float base = fog_density_godray(C);
  if (base <= 0.0f || t_end <= 0.0f) return V3s(0.0f);
  float phase = phase_blended(dot3(rd, SUN_DIR));
  float mu = dot3(rd, SUN_DIR);
  float view_gate = 0.10f + 0.90f * smoothstepf(0.0f, 0.20f, mu);
  uint32_t steps = godray_steps_for_tend(t_end);
  float dt = t_end / (float)steps;
  float a_ts = 1.0f - vx_exp(-dt * 3.0f);
  float a_shaft = 1.0f - vx_exp(-dt * 5.0f);
  float ts = 1.0f, sh = 0.0f;
  v3 sum = V3s(0.0f);
  float k = FOG_HEIGHT_FALLOFF, y0 = ro.y, dy = rd.y;
  float f = clampf(0.5f + j_phase, 0.0f, 1.0f);
  float one_m_f = 1.0f - f;
  float Tv_acc = 1.0f;
  bool horiz = std::fabs(dy) < 1e-4f;
  float tts_h = 1.0f, tte_h = 1.0f, inv_kdy = 0.0f, B = 0.0f, r_f = 1.0f, r_rem = 1.0f;
  if (horiz) {
    float a = vx_exp(-k * y0);
    float dens_const = base * a;
    tts_h = exp_neg_fast(dens_const * dt * f);
    tte_h = exp_neg_fast(dens_const * dt * one_m_f);
  } else {
    inv_kdy = 1.0f / (k * dy);
    B = vx_exp(-k * y0);
    r_f = vx_exp(-k * dy * dt * f);
    r_rem = vx_exp(-k * dy * dt * one_m_f);
  }
  float Ts_geom = 1.0f;
  for (uint32_t i = 0; i < steps; ++i) {
    float ti = ((float)i + 0.5f + j_phase) * dt;
    if (ti + j * dt > t_end) break;
    v3 p = ro + rd * ti;
    if (horiz) {
      Tv_acc *= tts_h;
    } else {
      float B_s = B * r_f;
      float d_od = base * (B - B_s) * inv_kdy;
      Tv_acc *= exp_neg_fast(mx(d_od, 0.0f));
    }
    float Tv = Tv_acc;
    if (Tv < GODRAY_TV_CUTOFF) break;
    if ((i & 3u) == 0u) Ts_geom = sun_transmittance_geom_only(C, p, SUN_DIR);
    float Tc = cloud_sun_transmittance_godray(C, p, SUN_DIR);
    float Tc_vol = mixf(1.0f, Tc, CLOUD_GODRAY_W);
    float Ts_clamped = clampf(Ts_geom, 0.0f, 1.0f);
    float Ts_soft = approx_pow_0p75(Ts_clamped);
    float haze_ramp = 1.0f - vx_exp(-ti / GODRAY_HAZE_NEAR_FADE);
    float height_term = mx(vx_exp(-GODRAY_SCATTER_HEIGHT_FALLOFF * p.y), GODRAY_SCATTER_MIN_FRAC);
    float dens = base * height_term;
    v3 common_factor = (SUN_COLOR * SUN_INTENSITY) * (dens * dt) * Tv * phase * Tc_vol * 0.70f;
    float ts_prev = ts;
    ts = mixf(ts, Ts_soft, a_ts);
    float dTs = mx(0.0f, ts_prev - Ts_soft);
    float shaft = smoothstepf(GODRAY_EDGE0, GODRAY_EDGE1, dTs);
    shaft *= (1.0f - Ts_soft);
    shaft = approx_pow_0p55(clampf(shaft, 0.0f, 1.0f));
    sh = mixf(sh, shaft, a_shaft);
    shaft = sh;
    float edge_boost = 1.0f + GODRAY_EDGE_ENERGY_BOOST * shaft;
    float shaft_sun_gate = smoothstepf(0.35f, 0.80f, ts);
    float haze = GODRAY_BASE_HAZE * haze_ramp * (ts * ts);
    float w = clampf(haze + (1.0f - haze) * (shaft * shaft_sun_gate), 0.0f, 1.0f);
    sum = sum + common_factor * view_gate * edge_boost * (ts * w);
    if (horiz) {
      Tv_acc *= tte_h;
    } else {
      float B_s = B * r_f;
      float B_e = B_s * r_rem;
      float d_od2 = base * (B_s - B_e) * inv_kdy;
      Tv_acc *= exp_neg_fast(mx(d_od2, 0.0f));
      B = B_e;
    }
  }
  v3 g = sum * GODRAY_ENERGY_BOOST;
  v3 knee = V3s(GODRAY_KNEE_INTEGRATE);
  return g / (g + knee);
}
static inline v3 godray_decompress(v3 c) {
  const float kk = 0.25f;
  v3 denom = mx3(V3s(1.0f) - c, V3s(1e-4f));
  return (kk * c) / denom;
}
static inline v3 godray_compress(v3 lin) { return lin / (lin + V3s(0.25f)); }

struct GodrayIO {
  uint32_t w, h;          // render dims (depth, godray, history)
  const float* depth;     // R
  const float* hist;      // RGBA
  float* out;             // RGBA
};
// compute_godray_pixel, godrays.wgsl:241-356 (main_godray, ray_main.wgsl:520-535)
static void godray_pixel(const Ctx& C, const GodrayIO& io, uint32_t gx, uint32_t gy) {
  int32_t fw = (int32_t)io.w, fh = (int32_t)io.h;
  v3 ro = cam_pos(C);
  int32_t ipx = (int32_t)gx, ipy = (int32_t)gy;
  v2 qpx = V2((float)gx, (float)gy);
  float fi = (float)(C.cam->frame_index & 1023u);
  v2 fj = V2(fi, fi * 1.371f);
  auto D = [&](int32_t x, int32_t y) {
    C.cnt->tex_rd += 4;
    return io.depth[(size_t)y * io.w + x];
  };
  float t_c = D(ipx, ipy);
  float t_l = D(clampi(ipx - 1, 0, fw - 1), ipy);
  float t_r = D(clampi(ipx + 1, 0, fw - 1), ipy);
  float t_u = D(ipx, clampi(ipy - 1, 0, fh - 1));
  float t_d = D(ipx, clampi(ipy + 1, 0, fh - 1));
  float dmin = mn(t_c, mn(mn(t_l, t_r), mn(t_u, t_d)));
  float dmax = mx(t_c, mx(mx(t_l, t_r), mx(t_u, t_d)));
  float span = (dmax - dmin) / mx(dmin, 1e-3f);
  float stable = 1.0f - smoothstepf(0.06f, 0.30f, span);
  float t_hist = dmin;
  float j4 = 0.20f * (hash12(qpx * J0_SCALE + fj) - 0.5f);
  float j_phase = (hash12(qpx * 0.91f + fj * 0.73f + V2(31.7f, 12.3f)) - 0.5f);
  const float qstep = 0.03f;
  float dq = (hash12(qpx * 1.37f + fj * 1.19f + V2(9.2f, 1.1f)) - 0.5f) * qstep;
  v2 j_rd = V2(hash12(qpx + fj + V2(1.7f, 9.2f)) - 0.5f, hash12(qpx + fj * 1.13f + V2(8.3f, 2.1f)) - 0.5f);
  float t_end = mn(std::floor((t_c + dq) / qstep) * qstep, GODRAY_MAX_DIST);
  bool fog_ok = fog_density_godray(C) > 0.0f;
  v2 px_center = V2((float)gx + 0.5f, (float)gy + 0.5f) + 0.35f * j_rd;
  v3 rd = ray_dir_from_pixel(C, px_center);
  v3 cur_lin = V3s(0.0f);
  if (fog_ok && t_end > 0.0f) cur_lin = godray_integrate_1(C, ro, rd, t_end, j4, j_phase);
  v3 p_ws = ro + rd * t_hist;
  v2 uv_prev = prev_uv_from_world(C, p_ws);
  v3 hist_lin = cur_lin;
  float hist_valid = 0.0f;
  v2 hdf = V2((float)io.w, (float)io.h);
  v2 uv_cur = V2((float)gx + 0.5f, (float)gy + 0.5f) / hdf;
  if (in_unit_square(uv_prev)) {
    int32_t sx = nearest_texel(uv_prev.x, io.w), sy = nearest_texel(uv_prev.y, io.h);
    const float* hh = io.hist + ((size_t)sy * io.w + sx) * 4;
    C.cnt->tex_rd += 16;
    float hist_depth = hh[3];
    hist_lin = godray_decompress(V3(hh[0], hh[1], hh[2]));
    float rel = std::fabs(hist_depth - t_hist) / mx(t_hist, 1e-3f);
    float depth_ok = 1.0f - smoothstepf(0.08f, 0.20f, rel);
    float depth_sane = hist_depth > 1e-3f ? 1.0f : 0.0f;
    float vel_px = length2((uv_prev - uv_cur) * hdf);
    float motion_ok = 1.0f - smoothstepf(0.75f, 2.50f, vel_px);
    hist_valid = depth_ok * depth_sane * motion_ok;
  }
  float delta_lin = length3(cur_lin - hist_lin);
  float energy = mx(length3(cur_lin), 1e-3f);
  float delta_rel = delta_lin / (0.05f + energy);
  float react = smoothstepf(0.10f, 0.45f, delta_rel);
  float clamp_scale = mixf(0.55f, 0.18f, stable);
  v3 clamp_w = mx3(cur_lin * clamp_scale, V3s(0.015f));
  v3 hist_clamped = clamp3(hist_lin, cur_lin - clamp_w, cur_lin + clamp_w);
  float hist_w_base = mixf(0.25f, 0.92f, stable);
  float hist_w = mixf(hist_w_base, 0.03f, react) * hist_valid;
  v3 out_lin = mix3(cur_lin, hist_clamped, hist_w);
  v3 oc = godray_compress(mx3(out_lin, V3s(0.0f)));
  store4(io.out, io.w, ipx, ipy, oc.x, oc.y, oc.z, t_hist);
  C.cnt->tex_wr += 16;
}

// ===========================================================================
// ray/composite.wgsl + main_composite (ray_main.wgsl:537-573)
// ===========================================================================
static inline v3 bright_extract_hue(v3 x, float thresh) {
  v3 lw = V3(0.2126f, 0.7152f, 0.0722f);
  float l = dot3(x, lw);
  float over = mx(l - thresh, 0.0f);
  float w = over / mx(l, 1e-4f);
  return x * w;
}
static inline v3 tonemap_filmic_hable(v3 x) {
  const float A = 0.22f, B = 0.30f, Cc = 0.10f, D = 0.20f, E = 0.01f, F = 0.30f;
  v3 num = x * (A * x + V3s(Cc * B)) + V3s(D * E);
  v3 den = x * (A * x + V3s(B)) + V3s(D * F);
  return (num / mx3(den, V3s(1e-6f))) - V3s(E / F);
}
static inline v3 tonemap_filmic_white_scale(v3 x, float white_point) {
  v3 w = tonemap_filmic_hable(V3s(white_point));
  v3 invw = V3s(1.0f) / mx3(w, V3s(1e-6f));
  return tonemap_filmic_hable(x) * invw;
}
struct CompositeIO {
  uint32_t rw, rh;            // render dims (color, godray, depth, local_hist)
  uint32_t pw, ph;            // present dims (output)
  const float* color;         // RGBA render
  const float* godray;        // RGBA render
  const float* depth;         // R render
  const float* local_hist;    // RGBA render
  float* out;                 // RGBA present
};
static void composite_pixel(const Ctx& C, const CompositeIO& io, uint32_t gx, uint32_t gy) {
  v2 pp = V2((float)gx + 0.5f, (float)gy + 0.5f);
  i2 ip_render = ip_render_from_present_px(C, pp);
  v2 px_render = px_render_from_present_px(C, pp);
  int32_t rwi = (int32_t)io.rw, rhi = (int32_t)io.rh;
  v2 rdf = V2((float)io.rw, (float)io.rh);
  i2 ipr{clampi(ip_render.x, 0, rwi - 1), clampi(ip_render.y, 0, rhi - 1)};
  auto COL = [&](int32_t x, int32_t y) {
    const float* p = io.color + ((size_t)y * io.rw + x) * 4;
    C.cnt->tex_rd += 16;
    return V3(p[0], p[1], p[2]);
  };
  auto DEP = [&](int32_t x, int32_t y) {
    C.cnt->tex_rd += 4;
    return io.depth[(size_t)y * io.rw + x];
  };
  v3 base = COL(ipr.x, ipr.y);
  // depth_full has the render dims.
  v2 fdf = rdf;
  v2 px_full = px_render * (fdf / rdf);
  i2 ipf{clampi(f2i(std::floor(px_full.x)), 0, rwi - 1), clampi(f2i(std::floor(px_full.y)), 0, rhi - 1)};
  v2 uv_full = (px_full + V2(0.5f, 0.5f)) / fdf;
  float d0 = DEP(ipf.x, ipf.y);
  v3 god_lin = V3s(0.0f);
  float god_far = 0.0f, god_scale = 0.0f;
  {
    v2 gdf = V2((float)io.rw, (float)io.rh);
    v2 du = V2(1.0f / gdf.x, 0.0f), dv = V2(0.0f, 1.0f / gdf.y);
    auto G = [&](v2 uv) {
      int32_t sx = nearest_texel(uv.x, io.rw), sy = nearest_texel(uv.y, io.rh);
      const float* p = io.godray + ((size_t)sy * io.rw + sx) * 4;
      C.cnt->tex_rd += 16;
      return godray_decompress(V3(p[0], p[1], p[2]));
    };
    v3 gC = G(uv_full), gE = G(uv_full + du), gW = G(uv_full - du), gN = G(uv_full + dv), gS = G(uv_full - dv);
    float dE = DEP(ipf.x + 1 < rwi - 1 ? ipf.x + 1 : rwi - 1, ipf.y);
    float dW = DEP(ipf.x - 1 > 0 ? ipf.x - 1 : 0, ipf.y);
    float dN = DEP(ipf.x, ipf.y + 1 < rhi - 1 ? ipf.y + 1 : rhi - 1);
    float dS = DEP(ipf.x, ipf.y - 1 > 0 ? ipf.y - 1 : 0);
    float tol = 0.02f + 0.06f * smoothstepf(10.0f, 80.0f, d0);
    float wE = 1.0f - smoothstepf(0.0f, tol, std::fabs(dE - d0));
    float wW = 1.0f - smoothstepf(0.0f, tol, std::fabs(dW - d0));
    float wN = 1.0f - smoothstepf(0.0f, tol, std::fabs(dN - d0));
    float wS = 1.0f - smoothstepf(0.0f, tol, std::fabs(dS - d0));
    float wsum = mx(wE + wW + wN + wS, 1e-4f);
    v3 blur = (gE * wE + gW * wW + gN * wN + gS * wS) / wsum;
    god_lin = mx3(gC + COMPOSITE_SHARPEN * (gC - blur), V3s(0.0f));
    god_lin = mx3(god_lin - V3s(GODRAY_BLACK_LEVEL), V3s(0.0f));
    god_lin = god_lin / (god_lin + V3s(GODRAY_KNEE_COMPOSITE));
    god_far = smoothstepf(GODRAY_FADE_NEAR, GODRAY_FADE_FAR, d0);
    god_scale = GODRAY_COMPOSITE_SCALE * mixf(1.0f, 0.25f, god_far);
  }
  v3 lum_w = V3(0.2126f, 0.7152f, 0.0722f);
  v3 sun_hue = normalize3(mx3(SUN_COLOR, V3s(1e-4f)));
  float g_lum = mx(dot3(god_lin, lum_w), 0.0f);
  v3 god_sun = sun_hue * g_lum;
  float hue_lock = 0.55f * god_far;
  god_lin = mix3(god_lin, god_sun, hue_lock);
  v3 warm = mix3(V3s(1.0f), V3(1.08f, 1.03f, 0.92f), god_far);
  god_lin = god_lin * warm;
  v3 hdr = mx3(base + god_scale * god_lin, V3s(0.0f));
  {
    const float bloom_thresh = 1.4f, bloom_k = 0.12f;
    float bloom_k_eff = bloom_k * mixf(1.0f, 0.0f, god_far);
    v3 b0 = bright_extract_hue(hdr, bloom_thresh);
    v3 hx1 = mx3(COL(clampi(ipr.x + 2, 0, rwi - 1), ipr.y), V3s(0.0f));
    v3 hx0 = mx3(COL(clampi(ipr.x - 2, 0, rwi - 1), ipr.y), V3s(0.0f));
    v3 hy1 = mx3(COL(ipr.x, clampi(ipr.y + 2, 0, rhi - 1)), V3s(0.0f));
    v3 hy0 = mx3(COL(ipr.x, clampi(ipr.y - 2, 0, rhi - 1)), V3s(0.0f));
    v3 bloom = (b0 + bright_extract_hue(hx1, bloom_thresh) + bright_extract_hue(hx0, bloom_thresh) +
                bright_extract_hue(hy1, bloom_thresh) + bright_extract_hue(hy0, bloom_thresh)) / 5.0f;
    v3 bloom_max = 0.35f * mx3(hdr, V3s(0.0f));
    hdr = hdr + bloom_k_eff * mn3(bloom, bloom_max);
  }
  float l_hdr = mx(dot3(hdr, lum_w), 1e-6f);
  v3 gray_h = V3s(l_hdr);
  float t_sat = smoothstepf(30.0f, 100.0f, d0);
  float sat_boost = 1.00f + 0.55f * t_sat;
  float hi = smoothstepf(1.6f, 6.0f, l_hdr);
  sat_boost = mixf(sat_boost, 1.0f, 0.55f * hi);
  hdr = mix3(gray_h, hdr, sat_boost);
  const float exposure = 1.10f, contrast = 1.05f, temp = 0.03f;
  hdr = hdr * exposure;
  hdr = hdr * V3(1.0f + temp, 1.0f, 1.0f - temp);
  v3 mid = V3s(0.18f);
  hdr = (hdr - mid) * contrast + mid;
  hdr = mx3(hdr, V3s(0.0f));
  v3 c = tonemap_filmic_white_scale(hdr * POST_EXPOSURE, 6.0f);
  c = clamp3s(c, 0.0f, 1.0f);
  c = V3(vx_pow(c.x, 0.98f), vx_pow(c.y, 0.98f), vx_pow(c.z, 0.98f));
  float fi = (float)(C.cam->frame_index & 255u);
  float n0 = hash12(px_render + V2(fi, 0.0f)) - 0.5f;
  float n1 = hash12(px_render * 0.73f + V2(0.0f, fi)) - 0.5f;
  float nn = 0.6f * n0 + 0.4f * n1;
  c = c + V3s(nn / 1536.0f);
  v3 cc = clamp3s(c, 0.0f, 1.0f);
  const float ig = 1.0f / 2.2f;
  v3 ldr = V3(vx_pow(cc.x, ig), vx_pow(cc.y, ig), vx_pow(cc.z, ig));
  // + local_hist sampled at the render UV (ray_main.wgsl:560-569).
  v2 inv_r = V2(1.0f / (float)io.rw, 1.0f / (float)io.rh);
  v2 uv_r = px_render * inv_r;
  int32_t lx = nearest_texel(uv_r.x, io.rw), ly = nearest_texel(uv_r.y, io.rh);
  const float* lp = io.local_hist + ((size_t)ly * io.rw + lx) * 4;
  C.cnt->tex_rd += 16;
  store4(io.out, io.pw, (int32_t)gx, (int32_t)gy, ldr.x + lp[0], ldr.y + lp[1], ldr.z + lp[2], 1.0f);
  C.cnt->tex_wr += 16;
}

}  // namespace vxo

// ===========================================================================
// C ABI for the tests (ctypes)
// ===========================================================================
extern "C" {

typedef struct vxo_scene {
  const vx_chunk_meta* chunks;
  uint64_t n_chunks;
  const vx_node* nodes;
  const vx_node_ropes* ropes;
  uint64_t n_nodes;
  const uint32_t* grid;
  uint64_t n_grid;
  const uint32_t* macro_words;
  uint64_t n_macro;
  const uint32_t* colinfo;
  uint64_t n_colinfo;
  const vx_clipmap* clip;
  const float* clip_height;
  uint64_t n_clip;
} vxo_scene;

typedef struct vxo_counts {
  uint64_t grid, meta, macro, node, rope, colinfo, clip, tex_rd, tex_wr, oob;
  uint64_t prof[5];  // PROFILE_VOXEL, GRASS, HDR, FOG, SHADOW (when cam.profile_flags & 1)
} vxo_counts;

}  // extern "C"

namespace {
vxo::Ctx make_ctx(const vxo_scene* s, const vx_camera* cam, vxo::Counts* cnt, bool godray_macro_compat) {
  vxo::Ctx C;
  C.cam = cam;
  C.chunks = s->chunks;
  C.nodes = s->nodes;
  C.chunk_grid = s->grid;
  C.macro_occ = s->macro_words;
  C.ropes = s->ropes;
  C.colinfo = s->colinfo;
  C.clip = s->clip;
  C.clip_height = s->clip_height;
  C.cnt = cnt;
  C.n_chunks = s->n_chunks;
  C.n_nodes = s->n_nodes;
  C.n_grid = s->n_grid;
  C.n_macro = s->n_macro;
  C.n_colinfo = s->n_colinfo;
  C.n_clip = s->n_clip;
  C.oob = 0;
  if (godray_macro_compat) {
    // The godray pass reads macro_occ through the binding that holds the rope
    // buffer (render/state/bindgroups.rs:142-151 vs common.wgsl:407-409).
    C.macro_occ = reinterpret_cast<const uint32_t*>(s->ropes);
    C.n_macro = s->n_nodes * 8;
  }
  return C;
}

template <class F>
void parallel_for(uint32_t n, int threads, F&& f) {
  if (threads <= 1 || n < 2) {
    for (uint32_t i = 0; i < n; ++i) f(i, 0);
    return;
  }
  std::atomic<uint32_t> next{0};
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; ++t)
    pool.emplace_back([&, t]() {
      for (;;) {
        uint32_t i = next.fetch_add(1);
        if (i >= n) return;
        f(i, t);
      }
    });
  for (auto& th : pool) th.join();
}

void add_counts(vxo_counts* dst, const vxo::Counts& c, uint64_t oob) {
  if (!dst) return;
  dst->grid += c.grid;
  dst->meta += c.meta;
  dst->macro += c.macro;
  dst->node += c.node;
  dst->rope += c.rope;
  dst->colinfo += c.colinfo;
  dst->clip += c.clip;
  dst->tex_rd += c.tex_rd;
  dst->tex_wr += c.tex_wr;
  dst->oob += oob;
  for (int i = 0; i < 5; ++i) dst->prof[i] += c.prof[i];
}

int hw_threads(int threads) {
  if (threads > 0) return threads;
  unsigned h = std::thread::hardware_concurrency();
  return h ? (int)h : 1;
}
}  // namespace

extern "C" {

// main_primary over the tiles [tile_begin, tile_end) of the (ceil(w/8) x
// ceil(h/8)) workgroup grid, tiles in row-major order.  tile_end = 0 means all.
int vxo_primary(const vxo_scene* scene, const vx_camera* cam, uint32_t w, uint32_t h, const float* primary_hist_in,
                const float* shadow_hist_in, float* color, float* depth, float* local, float* primary_hist_out,
                float* shadow_hist_out, vx_hit* hits, uint32_t tile_begin, uint32_t tile_end, int threads,
                vxo_counts* counts) {
  if (!scene || !cam || !w || !h || !primary_hist_in || !shadow_hist_in || !color || !depth || !local ||
      !primary_hist_out || !shadow_hist_out)
    return -1;
  const uint32_t tx = (w + 7) / 8, ty = (h + 7) / 8, nt = tx * ty;
  if (tile_end == 0 || tile_end > nt) tile_end = nt;
  if (tile_begin >= tile_end) return 0;
  vxo::PrimaryIO io{w, h, primary_hist_in, shadow_hist_in, color, depth, local, primary_hist_out, shadow_hist_out, hits};
  threads = hw_threads(threads);
  std::vector<vxo::Counts> cnt((size_t)threads);
  std::vector<uint64_t> oob((size_t)threads, 0);
  std::memset(cnt.data(), 0, sizeof(vxo::Counts) * cnt.size());
  parallel_for(tile_end - tile_begin, threads, [&](uint32_t i, int t) {
    vxo::Ctx C = make_ctx(scene, cam, &cnt[t], false);
    uint32_t tile = tile_begin + i;
    vxo::main_primary_tile(C, io, tile % tx, tile / tx);
    oob[t] += C.oob;
  });
  for (int t = 0; t < threads; ++t) add_counts(counts, cnt[t], oob[t]);
  return 0;
}

// main_local_taa, taa.wgsl:27-69 (nearest samples at pixel centres).
int vxo_local_taa(uint32_t w, uint32_t h, const float* local_in, const float* hist_in, float* hist_out) {
  using namespace vxo;
  if (!w || !h || !local_in || !hist_in || !hist_out) return -1;
  for (uint32_t y = 0; y < h; ++y)
    for (uint32_t x = 0; x < w; ++x) {
      v2 uv = V2((float)x + 0.5f, (float)y + 0.5f) / V2((float)w, (float)h);
      int32_t sx = nearest_texel(uv.x, w), sy = nearest_texel(uv.y, h);
      const float* c4 = local_in + ((size_t)sy * w + sx) * 4;
      const float* h4 = hist_in + ((size_t)sy * w + sx) * 4;
      v3 cur = V3(c4[0], c4[1], c4[2]);
      float w_in = c4[3];
      v3 hh = V3(h4[0], h4[1], h4[2]);
      bool hist_ok = !is_bad_vec3(hh);
      v3 hist = hist_ok ? hh : V3s(0.0f);
      float conf0 = hist_ok ? h4[3] : 0.0f;
      bool cur_ok = !is_bad_vec3(cur);
      bool has_sample = (w_in > 0.0f) && cur_ok;
      float k = has_sample ? LOCAL_TAA_ALPHA * clampf(w_in, 0.0f, 1.0f) : 0.0f;
      v3 out = mix3(hist, cur, k);
      float conf1 = has_sample ? mn(1.0f, conf0 + k) : conf0;
      store4(hist_out, w, (int32_t)x, (int32_t)y, out.x, out.y, out.z, conf1);
    }
  return 0;
}

// Ray queries: first voxel-geometry hit of each ray (rays as vx_ray, hits as
// vx_hit: kind 1 = hit, 0 = miss).
int vxo_query_rays(const vxo_scene* scene, const vx_camera* cam, const vx_ray* rays, uint32_t n, vx_hit* hits, int threads,
                   vxo_counts* counts) {
  if (!scene || !cam || (!rays && n) || (!hits && n)) return -1;
  threads = hw_threads(threads);
  std::vector<vxo::Counts> cnt((size_t)threads);
  std::vector<uint64_t> oob((size_t)threads, 0);
  std::memset(cnt.data(), 0, sizeof(vxo::Counts) * cnt.size());
  parallel_for(n, threads, [&](uint32_t i, int t) {
    vxo::Ctx C = make_ctx(scene, cam, &cnt[t], false);
    C.geom_only = true;
    const vx_ray& r = rays[i];
    vxo::HitGeom h = vxo::query_ray(C, vxo::V3(r.ox, r.oy, r.oz), vxo::V3(r.dx, r.dy, r.dz), r.tmax);
    vx_hit& o = hits[i];
    o.kind = h.hit ? 1u : 0u;
    o.slot = h.hit ? h.slot : VX_INVALID_U32;
    o.node = h.hit ? h.node : VX_INVALID_U32;
    o.voxel[0] = h.hit ? h.vox[0] : 0;
    o.voxel[1] = h.hit ? h.vox[1] : 0;
    o.voxel[2] = h.hit ? h.vox[2] : 0;
    o.mat = h.hit ? h.mat : VX_MAT_AIR;
    o.t = h.hit ? h.t : vxo::BIG_F32;
    oob[t] += C.oob;
  });
  for (int t = 0; t < threads; ++t) add_counts(counts, cnt[t], oob[t]);
  return 0;
}

int vxo_godray(const vxo_scene* scene, const vx_camera* cam, uint32_t w, uint32_t h, int macro_compat, const float* depth,
               const float* hist, float* out, int threads, vxo_counts* counts) {
  if (!scene || !cam || !w || !h || !depth || !hist || !out) return -1;
  vxo::GodrayIO io{w, h, depth, hist, out};
  threads = hw_threads(threads);
  std::vector<vxo::Counts> cnt((size_t)threads);
  std::vector<uint64_t> oob((size_t)threads, 0);
  std::memset(cnt.data(), 0, sizeof(vxo::Counts) * cnt.size());
  parallel_for(h, threads, [&](uint32_t y, int t) {
    vxo::Ctx C = make_ctx(scene, cam, &cnt[t], macro_compat != 0);
    for (uint32_t x = 0; x < w; ++x) vxo::godray_pixel(C, io, x, y);
    oob[t] += C.oob;
  });
  for (int t = 0; t < threads; ++t) add_counts(counts, cnt[t], oob[t]);
  return 0;
}

int vxo_composite(const vx_camera* cam, uint32_t rw, uint32_t rh, uint32_t pw, uint32_t ph, const float* color,
                  const float* godray, const float* depth, const float* local_hist, float* out, int threads) {
  if (!cam || !rw || !rh || !pw || !ph || !color || !godray || !depth || !local_hist || !out) return -1;
  vxo::CompositeIO io{rw, rh, pw, ph, color, godray, depth, local_hist, out};
  threads = hw_threads(threads);
  std::vector<vxo::Counts> cnt((size_t)threads);
  std::memset(cnt.data(), 0, sizeof(vxo::Counts) * cnt.size());
  vxo_scene empty;
  std::memset(&empty, 0, sizeof empty);
  parallel_for(ph, threads, [&](uint32_t y, int t) {
    vxo::Ctx C = make_ctx(&empty, cam, &cnt[t], false);
    for (uint32_t x = 0; x < pw; ++x) vxo::composite_pixel(C, io, x, y);
  });
  return 0;
}

// main_composite_taa, taa.wgsl:83-110.
int vxo_composite_taa(const vx_camera* cam, uint32_t rw, uint32_t rh, uint32_t pw, uint32_t ph, const float* composite_in,
                      const float* depth, const float* hist_in, float* output, float* hist_out) {
  using namespace vxo;
  if (!cam || !composite_in || !depth || !hist_in || !output || !hist_out) return -1;
  Counts cnt;
  std::memset(&cnt, 0, sizeof cnt);
  vxo_scene empty;
  std::memset(&empty, 0, sizeof empty);
  Ctx C = make_ctx(&empty, cam, &cnt, false);
  for (uint32_t y = 0; y < ph; ++y)
    for (uint32_t x = 0; x < pw; ++x) {
      v2 pp = V2((float)x + 0.5f, (float)y + 0.5f);
      i2 ipr = ip_render_from_present_px(C, pp);
      v2 pr = px_render_from_present_px(C, pp);
      const float* c4 = composite_in + ((size_t)y * pw + x) * 4;
      v3 taa = V3(c4[0], c4[1], c4[2]);
      if (cam->frame_index > 1u) {
        float d = depth[(size_t)ipr.y * rw + ipr.x];
        v3 rd = ray_dir_from_pixel(C, pr);
        v3 p_ws = cam_pos(C) + rd * d;
        v2 uvp = prev_uv_from_world(C, p_ws);
        if (in_unit_square(uvp)) {
          int32_t sx = nearest_texel(uvp.x, pw), sy = nearest_texel(uvp.y, ph);
          const float* h4 = hist_in + ((size_t)sy * pw + sx) * 4;
          taa = mix3(V3(h4[0], h4[1], h4[2]), taa, COMPOSITE_TAA_ALPHA);
        }
      }
      store4(output, pw, (int32_t)x, (int32_t)y, taa.x, taa.y, taa.z, 1.0f);
      store4(hist_out, pw, (int32_t)x, (int32_t)y, taa.x, taa.y, taa.z, 1.0f);
    }
  return 0;
}

// Scalar math entry points, for the accuracy tests of the shared method.
float vxo_sin(float x) { return vxo::vx_sin(x); }
float vxo_cos(float x) { return vxo::vx_cos(x); }
float vxo_exp(float x) { return vxo::vx_exp(x); }
float vxo_log(float x) { return vxo::vx_log(x); }
float vxo_pow(float x, float y) { return vxo::vx_pow(x, y); }
uint32_t vxo_hash_u32(uint32_t x) { return vxo::hash_u32(x); }
float vxo_hash12(float x, float y) { return vxo::hash12(vxo::V2(x, y)); }
float vxo_value_noise(float x, float y) { return vxo::value_noise(vxo::V2(x, y)); }
float vxo_fbm(float x, float y) { return vxo::fbm(vxo::V2(x, y)); }
void vxo_tonemap(const float* in3, float* out3) {
  vxo::v3 r = vxo::tonemap_filmic_white_scale(vxo::V3(in3[0], in3[1], in3[2]), 6.0f);
  out3[0] = r.x;
  out3[1] = r.y;
  out3[2] = r.z;
}
float vxo_phase_mie(float c) { return vxo::phase_mie(c); }
void vxo_intersect_aabb(const float* ro, const float* rd, const float* bmin, const float* bmax, float* out2) {
  vxo::v2 r = vxo::intersect_aabb(vxo::V3(ro[0], ro[1], ro[2]), vxo::V3(rd[0], rd[1], rd[2]), vxo::V3(bmin[0], bmin[1], bmin[2]),
                                  vxo::V3(bmax[0], bmax[1], bmax[2]));
  out2[0] = r.x;
  out2[1] = r.y;
}
float vxo_fog_od(float base, const float* ro, const float* rd, float t) {
  return vxo::fog_optical_depth_with_base(base, vxo::V3(ro[0], ro[1], ro[2]), vxo::V3(rd[0], rd[1], rd[2]), t);
}

// The tile candidate gather of main_primary (ray_main.wgsl:179-197) alone:
// per 8x8 tile, the raw candidate count the two probe rays produce (the
// atomic counter value, which may exceed MAX_TILE_CHUNKS) — for tests that
// must know whether a frame reaches the 24-candidate truncation.
int vxo_tile_raw_counts(const vxo_scene* scene, const vx_camera* cam, uint32_t w, uint32_t h, uint32_t* raw) {
  if (!scene || !cam || !w || !h || !raw) return -1;
  const uint32_t tx = (w + 7) / 8, ty = (h + 7) / 8;
  vxo::Counts cnt;
  std::memset(&cnt, 0, sizeof(cnt));
  for (uint32_t t = 0; t < tx * ty; ++t) {
    vxo::Ctx C = make_ctx(scene, cam, &cnt, false);
    vxo::TileList L;
    L.count = 0;
    if (cam->chunk_count != 0u) {
      const uint32_t wgx = t % tx, wgy = t / tx;
      vxo::v2 base = vxo::V2((float)(wgx * vxo::TILE_SIZE), (float)(wgy * vxo::TILE_SIZE));
      vxo::tile_append_candidates_for_ray(C, L, vxo::cam_pos(C), vxo::ray_dir_from_pixel(C, base + vxo::V2(4.5f, 4.5f)),
                                          0.0f, vxo::FOG_MAX_DIST);
      vxo::tile_append_candidates_for_ray(C, L, vxo::cam_pos(C), vxo::ray_dir_from_pixel(C, base + vxo::V2(1.5f, 1.5f)),
                                          0.0f, vxo::FOG_MAX_DIST);
    }
    raw[t] = L.count;
  }
  return 0;
}

}  // extern "C"
