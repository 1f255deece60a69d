// vxo_core.h — oracle restatement of the reference WGSL shared code:
// common.wgsl, ray/{clouds,phase,sky,fog,aabb,wind,svo_query,leaves,grass}.wgsl.
//
// TEST INFRASTRUCTURE ONLY (the checker of the HIP passes; never shipped).
// Every function names the WGSL function and file:line it restates; paths are
// relative to src/shaders/ of the reference.  Arithmetic follows the WGSL
// expression order operand for operand (IEEE binary32, no contraction).
#pragma once
#include "vxo_math.h"
#include "../include/voxhip.h"

namespace vxo {

// ---- constants: common.wgsl:16-332 ---------------------------------------
constexpr uint32_t LEAF_U32 = 0xFFFFFFFFu;
constexpr uint32_t MAT_AIR = 0, MAT_GRASS = 1, MAT_DIRT = 2, MAT_STONE = 3, MAT_WOOD = 4, MAT_LEAF = 5, MAT_LIGHT = 6;
static const v3 SUN_DIR = {0.61237244f, 0.5f, 0.61237244f};
static const v3 SUN_COLOR = {1.0f, 0.94f, 0.72f};
constexpr float SUN_INTENSITY = 5.0f;
constexpr float SHADOW_BIAS = 2e-4f;
constexpr uint32_t VSM_STEPS = 16;
constexpr float LEAF_LIGHT_TRANSMIT = 0.50f, GRASS_LIGHT_TRANSMIT = 0.70f, MIN_TRANS = 0.03f;
constexpr float GODRAY_ENERGY_BOOST = 8.0f, GODRAY_KNEE_INTEGRATE = 0.35f, GODRAY_COMPOSITE_SCALE = 6.5f;
constexpr float GODRAY_FADE_NEAR = 40.0f, GODRAY_FADE_FAR = 80.0f, GODRAY_KNEE_COMPOSITE = 0.25f;
constexpr float FOG_HEIGHT_FALLOFF = 0.18f, FOG_MAX_DIST = 120.0f, FOG_PRIMARY_SCALE = 0.02f, FOG_GODRAY_SCALE = 1.0f;
static const v3 FOG_COLOR_GROUND = {0.40f, 0.42f, 0.45f};
constexpr float FOG_COLOR_SKY_BLEND = 0.10f;
constexpr float GODRAY_MAX_DIST = 100.0f, GODRAY_SCATTER_HEIGHT_FALLOFF = 0.04f, GODRAY_SCATTER_MIN_FRAC = 0.35f;
constexpr float GODRAY_BLACK_LEVEL = 0.004f, GODRAY_EDGE0 = 0.004f, GODRAY_EDGE1 = 0.035f;
constexpr float GODRAY_BASE_HAZE = 0.02f, GODRAY_HAZE_NEAR_FADE = 18.0f, GODRAY_TV_CUTOFF = 0.02f;
constexpr uint32_t GODRAY_STEPS_FAST = 24, GODRAY_MIN_STEPS = 8, GODRAY_STEP_Q = 4;
constexpr float GODRAY_STEPS_PER_METER = 1.25f, GODRAY_EDGE_ENERGY_BOOST = 2.5f;
constexpr float INV_4PI = 0.0795774715f, PHASE_G = 0.15f, PHASE_MIE_W = 0.20f;
constexpr float CLOUD_BASE_H = 170.0f, CLOUD_TOP_H = 250.0f, CLOUD_UV_SCALE = 0.0016f;
static const v2 CLOUD_WIND = {0.020f, 0.012f};
constexpr float CLOUD_COVERAGE = 0.40f, CLOUD_SOFTNESS = 0.08f, CLOUD_DENSITY = 0.058f;
constexpr float CLOUD_PUFF_POW = 1.8f, CLOUD_DETAIL_W = 0.35f;
constexpr uint32_t CLOUD_STEPS_VIEW = 8, CLOUD_STEPS_LIGHT = 4;
constexpr float CLOUD_HORIZON_Y0 = 0.02f, CLOUD_HORIZON_Y1 = 0.25f;
static const v3 CLOUD_BASE_COL = {0.72f, 0.74f, 0.76f};
constexpr float CLOUD_SILVER_POW = 8.0f, CLOUD_SILVER_STR = 0.6f;
constexpr float CLOUD_SHADOW_ABSORB = 6.0f, CLOUD_SHADOW_STRENGTH = 0.45f;
constexpr float CLOUD_SUN_DISC_DIM_POW = 8.0f, CLOUD_SUN_DISC_DIM_FLOOR = 0.0f;
constexpr float CLOUD_GODRAY_W = 0.50f, SKY_EXPOSURE = 0.40f;
constexpr float WIND_CELL_FREQ = 2.5f;
static const v2 WIND_DIR_XZ = {0.9f, 0.4f};
constexpr float WIND_RAMP_Y0 = 2.0f, WIND_RAMP_Y1 = 14.0f;
constexpr float WIND_GUST_TIME_FREQ = 0.9f, WIND_FLUTTER_TIME_FREQ = 4.2f;
static const v2 WIND_GUST_XZ_FREQ = {0.35f, 0.22f};
static const v2 WIND_FLUTTER_XZ_FREQ = {1.7f, 1.1f};
constexpr float WIND_GUST_WEIGHT = 0.75f, WIND_FLUTTER_WEIGHT = 0.25f;
constexpr float WIND_VERTICAL_SCALE = 0.25f, LEAF_VERTICAL_REDUCE = 0.15f;
constexpr float LEAF_OFFSET_AMP = 0.75f, LEAF_OFFSET_MAX_FRAC = 0.75f;
constexpr float LEAF_LOD_DISP_START = 25.0f, LEAF_LOD_DISP_END = 70.0f;
static const v3 WIND_PHASE_OFF_1 = {19.0f, 7.0f, 11.0f};
constexpr float TAU = 6.28318530718f;
constexpr float PRIMARY_NUDGE_VOXEL_FRAC = 1e-3f;
constexpr float PRIMARY_HIT_MARGIN = 0.08f, PRIMARY_HIT_WINDOW = 0.60f;
constexpr float PRIMARY_HIT_DEPTH_REL0 = 0.04f, PRIMARY_HIT_DEPTH_REL1 = 0.12f;
constexpr float SHADOW_TAA_ALPHA = 0.20f;
constexpr uint32_t SHADOW_SUBSAMPLE_MASK = 7;
constexpr float PRIMARY_HIT_MOTION_PX0 = 0.50f, PRIMARY_HIT_MOTION_PX1 = 1.75f;
constexpr float J0_SCALE = 1.31f;
constexpr float COMPOSITE_SHARPEN = 0.1f, POST_EXPOSURE = 0.15f;
constexpr uint32_t CLIP_LEVELS_MAX = 16, HF_MAX_STEPS = 48, HF_BISECT = 3;
constexpr float HF_DT_MAX = 48.0f;
constexpr uint32_t LIGHT_MAX_DIST_VOX = 16, LIGHT_RAYS = 8, LIGHT_EARLY_HITS = 4;
constexpr float LIGHT_SOFT_RADIUS_VOX = 3.0f, LIGHT_NEAR_CLAMP_VOX = 1.25f, LIGHT_RANGE_VOX = 80.0f;
constexpr float LIGHT_WRAP = 0.15f, LIGHT_DIRECT_GAIN = 1.00f, LIGHT_INDIRECT_GAIN = 0.65f;
constexpr float VOXEL_AO_MAX_DIST = 40.0f, LOCAL_LIGHT_MAX_DIST = 50.0f, FAR_SHADING_DIST = 64.0f;
constexpr float LOCAL_TAA_ALPHA = 0.80f, COMPOSITE_TAA_ALPHA = 0.45f;
constexpr float GRASS_LAYER_HEIGHT_VOX = 1.55f;
constexpr uint32_t GRASS_BLADE_COUNT = 3, GRASS_TRACE_STEPS = 7;
constexpr float GRASS_HIT_EPS_VOX = 0.008f, GRASS_STEP_MIN_VOX = 0.02f, GRASS_VOXEL_SEGS = 6.0f;
constexpr float GRASS_VOXEL_THICKNESS_VOX = 0.15f, GRASS_VOXEL_TAPER = 0.25f, GRASS_OVERHANG_VOX = 0.6f;
constexpr float GRASS_LOD_MID_START = 18.0f, GRASS_LOD_FAR_START = 45.0f;
constexpr uint32_t GRASS_BLADE_COUNT_MID = 2, GRASS_BLADE_COUNT_FAR = 1;
constexpr uint32_t GRASS_SEGS_MID = 4, GRASS_SEGS_FAR = 2, GRASS_TRACE_STEPS_MID = 5, GRASS_TRACE_STEPS_FAR = 3;
constexpr float ALBEDO_VAR_GAIN = 3.5f;
constexpr uint32_t MACRO_DIM = 8, TILE_SIZE = 8, MAX_TILE_CHUNKS = 192, PRIMARY_MAX_TILE_CHUNKS = 24;
constexpr float U32_TO_F01 = 1.0f / 4294967296.0f;

// ---- algorithmic load counters (the per-ray byte model) -------------------
struct Counts {
  uint64_t grid, meta, macro, node, rope, colinfo, clip;  // loads by array
  uint64_t tex_rd, tex_wr;                                 // history/texture bytes
  // The primary pass's profile counters (ray_main.wgsl:32-47: PROFILE_VOXEL,
  // GRASS, HDR, FOG, SHADOW), raised only when cam.profile_flags & 1.
  uint64_t prof[5];
};

// ---- bindings -------------------------------------------------------------
struct Ctx {
  const vx_camera* cam;
  const vx_chunk_meta* chunks;
  const vx_node* nodes;
  const uint32_t* chunk_grid;
  const uint32_t* macro_occ;  // the godray pass may alias this to the rope words
  const vx_node_ropes* ropes;
  const uint32_t* colinfo;
  const vx_clipmap* clip;
  const float* clip_height;
  Counts* cnt;
  // Bounds of each array, in elements (reads past them return zero, as a
  // robust-access implementation would; tests assert none happen).
  uint64_t n_chunks, n_nodes, n_grid, n_macro, n_colinfo, n_clip;
  mutable uint64_t oob;
  // Ray queries trace voxel geometry only: no grass blades, no leaf displacement.
  bool geom_only = false;
};

static inline vx_chunk_meta ld_chunk(const Ctx& C, uint32_t slot) {
  C.cnt->meta++;
  if (slot >= C.n_chunks) { C.oob++; vx_chunk_meta z; std::memset(&z, 0, sizeof z); return z; }
  return C.chunks[slot];
}
static inline vx_node ld_node(const Ctx& C, uint64_t idx) {
  C.cnt->node++;
  if (idx >= C.n_nodes) { C.oob++; vx_node z; std::memset(&z, 0, sizeof z); return z; }
  return C.nodes[idx];
}
static inline vx_node_ropes ld_ropes(const Ctx& C, uint64_t idx) {
  C.cnt->rope++;
  if (idx >= C.n_nodes) { C.oob++; vx_node_ropes z; std::memset(&z, 0, sizeof z); return z; }
  return C.ropes[idx];
}
static inline uint32_t ld_grid(const Ctx& C, uint64_t idx) {
  C.cnt->grid++;
  if (idx >= C.n_grid) { C.oob++; return 0; }
  return C.chunk_grid[idx];
}
static inline uint32_t ld_macro(const Ctx& C, uint64_t idx) {
  C.cnt->macro++;
  if (idx >= C.n_macro) { C.oob++; return 0; }
  return C.macro_occ[idx];
}
static inline uint32_t ld_colinfo(const Ctx& C, uint64_t idx) {
  C.cnt->colinfo++;
  if (idx >= C.n_colinfo) { C.oob++; return 0; }
  return C.colinfo[idx];
}
static inline float ld_clip(const Ctx& C, uint32_t layer, int32_t sx, int32_t sz) {
  C.cnt->clip++;
  uint64_t idx = ((uint64_t)layer * C.clip->res + (uint64_t)sz) * C.clip->res + (uint64_t)sx;
  if (idx >= C.n_clip) { C.oob++; return 0.0f; }
  return C.clip_height[idx];
}

static inline v3 cam_pos(const Ctx& C) { return V3(C.cam->cam_pos[0], C.cam->cam_pos[1], C.cam->cam_pos[2]); }
static inline float vp_x(const Ctx& C) { return C.cam->voxel_params[0]; }

// ===========================================================================
// common.wgsl
// ===========================================================================
// macro_bit_index, common.wgsl:431-433
static inline uint32_t macro_bit_index(uint32_t mx_, uint32_t my_, uint32_t mz_) {
  return mx_ + MACRO_DIM * (my_ + MACRO_DIM * mz_);
}
// macro_test, common.wgsl:435-440
static inline bool macro_test(const Ctx& C, uint32_t macro_base, uint32_t bit) {
  uint32_t w = bit >> 5u, b = bit & 31u;
  uint32_t word = ld_macro(C, (uint64_t)macro_base + w);
  return (word & (1u << b)) != 0u;
}
// safe_inv, common.wgsl:442-444
static inline float safe_inv(float x) { return std::fabs(x) < EPS_INV ? BIG_F32 : 1.0f / x; }
static inline v3 safe_inv3(v3 d) { return V3(safe_inv(d.x), safe_inv(d.y), safe_inv(d.z)); }
// safe_normalize, common.wgsl:446-453
static inline v3 safe_normalize(v3 v) {
  float l2 = dot3(v, v);
  if (l2 <= 1e-12f) return V3(0.0f, 1.0f, 0.0f);
  return v * inverse_sqrt(l2);
}
// ray_dir_from_pixel, common.wgsl:456-459
static inline v3 ray_dir_from_pixel(const Ctx& C, v2 px) {
  const vx_camera* c = C.cam;
  v3 r00 = V3(c->ray00[0], c->ray00[1], c->ray00[2]);
  v3 rdx = V3(c->ray_dx[0], c->ray_dx[1], c->ray_dx[2]);
  v3 rdy = V3(c->ray_dy[0], c->ray_dy[1], c->ray_dy[2]);
  v3 d = r00 + px.x * rdx + px.y * rdy;
  return normalize3(d);
}
// mat4 (column-major) * vec4: sum over columns in order.
static inline v4 mat4_mul(const float* m, v4 v) {
  v4 r;
  r.x = ((m[0] * v.x + m[4] * v.y) + m[8] * v.z) + m[12] * v.w;
  r.y = ((m[1] * v.x + m[5] * v.y) + m[9] * v.z) + m[13] * v.w;
  r.z = ((m[2] * v.x + m[6] * v.y) + m[10] * v.z) + m[14] * v.w;
  r.w = ((m[3] * v.x + m[7] * v.y) + m[11] * v.z) + m[15] * v.w;
  return r;
}
// prev_uv_from_world, common.wgsl:461-467
static inline v2 prev_uv_from_world(const Ctx& C, v3 p) {
  v4 clip = mat4_mul(C.cam->prev_view_proj, v4{p.x, p.y, p.z, 1.0f});
  float invw = 1.0f / mx(clip.w, 1e-6f);
  v2 ndc = V2(clip.x * invw, clip.y * invw);
  v2 uv = ndc * 0.5f + 0.5f;
  return V2(uv.x, 1.0f - uv.y);
}
// in_unit_square, common.wgsl:469-471
static inline bool in_unit_square(v2 uv) { return uv.x >= 0.0f && uv.y >= 0.0f && uv.x <= 1.0f && uv.y <= 1.0f; }
// intersect_aabb, common.wgsl:473-510
static inline v2 intersect_aabb(v3 ro, v3 rd, v3 bmin, v3 bmax) {
  const float eps = 1e-8f;
  float te = -1e30f, tx = 1e30f;
  if (std::fabs(rd.x) < eps) {
    if (ro.x < bmin.x || ro.x > bmax.x) return V2(1.0f, 0.0f);
  } else {
    float inv = 1.0f / rd.x;
    float t0 = (bmin.x - ro.x) * inv, t1 = (bmax.x - ro.x) * inv;
    te = mx(te, mn(t0, t1));
    tx = mn(tx, mx(t0, t1));
  }
  if (std::fabs(rd.y) < eps) {
    if (ro.y < bmin.y || ro.y > bmax.y) return V2(1.0f, 0.0f);
  } else {
    float inv = 1.0f / rd.y;
    float t0 = (bmin.y - ro.y) * inv, t1 = (bmax.y - ro.y) * inv;
    te = mx(te, mn(t0, t1));
    tx = mn(tx, mx(t0, t1));
  }
  if (std::fabs(rd.z) < eps) {
    if (ro.z < bmin.z || ro.z > bmax.z) return V2(1.0f, 0.0f);
  } else {
    float inv = 1.0f / rd.z;
    float t0 = (bmin.z - ro.z) * inv, t1 = (bmax.z - ro.z) * inv;
    te = mx(te, mn(t0, t1));
    tx = mn(tx, mx(t0, t1));
  }
  return V2(te, tx);
}
// child_rank, common.wgsl:512-516
static inline uint32_t child_rank(uint32_t mask, uint32_t ci) {
  uint32_t bit = 1u << ci;
  return (uint32_t)__builtin_popcount(mask & (bit - 1u));
}
// grid_lookup_slot, common.wgsl:518-541
static inline uint32_t grid_lookup_slot(const Ctx& C, int32_t cx, int32_t cy, int32_t cz) {
  const vx_camera* c = C.cam;
  int32_t ix = cx - c->grid_origin_chunk[0], iy = cy - c->grid_origin_chunk[1], iz = cz - c->grid_origin_chunk[2];
  if (ix < 0 || iy < 0 || iz < 0) return INVALID_U32;
  uint32_t nx = c->grid_dims[0], ny = c->grid_dims[1], nz = c->grid_dims[2];
  if ((uint32_t)ix >= nx || (uint32_t)iy >= ny || (uint32_t)iz >= nz) return INVALID_U32;
  uint32_t idx = ((uint32_t)iz * ny * nx) + ((uint32_t)iy * nx) + (uint32_t)ix;
  return ld_grid(C, idx);
}
struct i3 { int32_t x, y, z; };
// chunk_coord_from_pos, common.wgsl:543-554
static inline i3 chunk_coord_from_pos(v3 p, float cs) {
  float eps = 1e-6f * cs;
  float px = p.x - signf(p.x) * eps, py = p.y - signf(p.y) * eps, pz = p.z - signf(p.z) * eps;
  return i3{f2i(std::floor(px / cs)), f2i(std::floor(py / cs)), f2i(std::floor(pz / cs))};
}
// chunk_max_depth, common.wgsl:569-572
static inline uint32_t chunk_max_depth(const Ctx& C) { return 31u - (uint32_t)__builtin_clz(C.cam->chunk_size); }
// hash_u32 / hash2_u32 / hash12 / hash3_u32 / hash31, common.wgsl:597-631
static inline uint32_t hash_u32(uint32_t v) {
  v ^= v >> 16u;
  v *= 0x7feb352du;
  v ^= v >> 15u;
  v *= 0x846ca68bu;
  v ^= v >> 16u;
  return v;
}
static inline uint32_t hash2_u32(uint32_t x, uint32_t y) { return hash_u32(x * 0x8da6b343u ^ y * 0xd8163841u); }
static inline float hash12(v2 p) {
  uint32_t ix = (uint32_t)f2i(std::floor(p.x)), iy = (uint32_t)f2i(std::floor(p.y));
  return (float)hash2_u32(ix, iy) * U32_TO_F01;
}
static inline uint32_t hash3_u32(uint32_t x, uint32_t y, uint32_t z) {
  return hash_u32(x * 0x8da6b343u ^ y * 0xd8163841u ^ z * 0xcb1ab31fu);
}
static inline float hash31(v3 p) {
  uint32_t ix = (uint32_t)f2i(std::floor(p.x)), iy = (uint32_t)f2i(std::floor(p.y)), iz = (uint32_t)f2i(std::floor(p.z));
  return (float)hash3_u32(ix, iy, iz) * U32_TO_F01;
}
// value_noise, common.wgsl:633-647
static inline float value_noise(v2 p) {
  v2 i = floor2(p);
  v2 f = fract2(p);
  float a = hash12(i);
  float b = hash12(i + V2(1.0f, 0.0f));
  float c = hash12(i + V2(0.0f, 1.0f));
  float d = hash12(i + V2(1.0f, 1.0f));
  v2 u = f * f * V2(3.0f - 2.0f * f.x, 3.0f - 2.0f * f.y);
  float x1 = mixf(a, b, u.x);
  float x2 = mixf(c, d, u.x);
  return mixf(x1, x2, u.y);
}
// fbm, common.wgsl:649-662 (rot = mat2x2(0.8, -0.6, 0.6, 0.8), column-major)
static inline float fbm(v2 p) {
  v2 x = p;
  float sum = 0.0f, amp = 0.5f;
  for (uint32_t i = 0; i < 5u; ++i) {
    sum += amp * value_noise(x);
    v2 r = V2(0.8f * x.x + 0.6f * x.y, -0.6f * x.x + 0.8f * x.y);
    x = r * 2.0f + V2(17.0f, 9.0f);
    amp *= 0.5f;
  }
  return sum;
}
// render/present mapping, common.wgsl:664-690
static inline v2 render_dims_f(const Ctx& C) { return V2((float)C.cam->render_present_px[0], (float)C.cam->render_present_px[1]); }
static inline v2 present_dims_f(const Ctx& C) { return V2((float)C.cam->render_present_px[2], (float)C.cam->render_present_px[3]); }
static inline v2 px_render_from_present_px(const Ctx& C, v2 pp) {
  v2 rdim = render_dims_f(C);
  v2 uv = pp / present_dims_f(C);
  return uv * rdim;
}
struct i2 { int32_t x, y; };
static inline int32_t clampi(int32_t v, int32_t lo, int32_t hi) { return v < lo ? lo : (v > hi ? hi : v); }
static inline i2 ip_render_from_present_px(const Ctx& C, v2 pp) {
  v2 rdim = render_dims_f(C);
  v2 pr = px_render_from_present_px(C, pp);
  return i2{clampi(f2i(std::floor(pr.x)), 0, f2i(rdim.x) - 1), clampi(f2i(std::floor(pr.y)), 0, f2i(rdim.y) - 1)};
}
// is_bad_vec3, common.wgsl:692-706
static inline bool is_bad_vec3(v3 v) {
  auto bad = [](float x) { return x != x || std::fabs(x) > 1.0e30f; };
  return bad(v.x) || bad(v.y) || bad(v.z);
}
// ray_eps_vec, common.wgsl:708-716
static inline v3 ray_eps_vec(v3 rd, float eps) {
  return V3(rd.x >= 0.0f ? eps : -eps, rd.y >= 0.0f ? eps : -eps, rd.z >= 0.0f ? eps : -eps);
}

// ===========================================================================
// ray/clouds.wgsl, ray/phase.wgsl, ray/sky.wgsl
// ===========================================================================
// cloud_height01, clouds.wgsl:5-15
static inline float cloud_height01(float y) {
  float t = clampf((y - CLOUD_BASE_H) / mx(CLOUD_TOP_H - CLOUD_BASE_H, 1e-3f), 0.0f, 1.0f);
  float edge_in = smoothstepf(0.00f, 0.12f, t);
  float edge_out = 1.0f - smoothstepf(0.78f, 1.00f, t);
  return edge_in * edge_out;
}
// cloud_coverage_field, clouds.wgsl:18-24
static inline float cloud_coverage_field(v2 xz, float time_s) {
  v2 uv = xz * CLOUD_UV_SCALE + CLOUD_WIND * time_s;
  float n = fbm(uv);
  float n2 = fbm(uv * 2.3f + V2(13.2f, 7.1f));
  return 0.65f * n + 0.35f * n2;
}
// cloud_coverage_at_xz, clouds.wgsl:27-31
static inline float cloud_coverage_at_xz(v2 xz, float time_s) {
  float field = cloud_coverage_field(xz, time_s);
  return smoothstepf(CLOUD_COVERAGE, CLOUD_COVERAGE + CLOUD_SOFTNESS, field);
}
// cloud_density, clouds.wgsl:34-66
static inline float cloud_density(v3 p, float time_s) {
  if (p.y <= CLOUD_BASE_H || p.y >= CLOUD_TOP_H) return 0.0f;
  float h = cloud_height01(p.y);
  if (h <= 0.0f) return 0.0f;
  float cov = cloud_coverage_at_xz(V2(p.x, p.z), time_s);
  if (cov <= 0.001f) return 0.0f;
  v2 uv = V2(p.x, p.z) * CLOUD_UV_SCALE + CLOUD_WIND * time_s;
  uv = uv + 0.003f * p.y;
  float base = fbm(uv * 1.2f);
  float det = fbm(uv * 3.7f + V2(0.11f * p.y, -0.07f * p.y));
  float d = mixf(base, det, CLOUD_DETAIL_W);
  d = smoothstepf(0.45f, 0.95f, d);
  d = vx_pow(d, CLOUD_PUFF_POW);
  d *= cov;
  d *= h;
  return clampf(d, 0.0f, 1.0f);
}
// cloud_light_transmittance, clouds.wgsl:69-91
static inline float cloud_light_transmittance(v3 p, float time_s) {
  if (SUN_DIR.y <= 0.01f) return 1.0f;
  float tmax = (CLOUD_TOP_H - p.y) / mx(SUN_DIR.y, 1e-3f);
  if (tmax <= 0.0f) return 1.0f;
  float dt = tmax / (float)CLOUD_STEPS_LIGHT;
  float od = 0.0f;
  for (uint32_t i = 0; i < CLOUD_STEPS_LIGHT; ++i) {
    float ti = ((float)i + 0.5f) * dt;
    v3 ps = p + SUN_DIR * ti;
    od += cloud_density(ps, time_s) * dt;
  }
  return vx_exp(-CLOUD_DENSITY * od * CLOUD_SHADOW_ABSORB);
}
// cloud_sun_transmittance_fast, clouds.wgsl:135-166
static inline float cloud_sun_transmittance_fast(const Ctx& C, v3 p, v3 sun_dir) {
  if (sun_dir.y <= 0.01f) return 1.0f;
  float t0 = (CLOUD_BASE_H - p.y) / sun_dir.y, t1 = (CLOUD_TOP_H - p.y) / sun_dir.y;
  float ta = mn(t0, t1), tb = mx(t0, t1);
  float t_start = mx(ta, 0.0f), t_end = mx(tb, 0.0f);
  if (t_end <= t_start) return 1.0f;
  float time_s = C.cam->voxel_params[1];
  float dt = (t_end - t_start) / 2.0f;
  float od = 0.0f;
  for (uint32_t i = 0; i < 2u; ++i) {
    float ti = t_start + ((float)i + 0.5f) * dt;
    v3 ps = p + sun_dir * ti;
    od += cloud_density(ps, time_s) * dt;
  }
  return vx_exp(-CLOUD_DENSITY * od * CLOUD_SHADOW_ABSORB);
}
// cloud_density_godray, clouds.wgsl:168-183
static inline float cloud_density_godray(v3 p, float time_s) {
  if (p.y <= CLOUD_BASE_H || p.y >= CLOUD_TOP_H) return 0.0f;
  float h = cloud_height01(p.y);
  if (h <= 0.0f) return 0.0f;
  v2 uv = V2(p.x, p.z) * CLOUD_UV_SCALE + CLOUD_WIND * time_s;
  float n0 = value_noise(uv);
  float n1 = value_noise(uv * 2.0f + V2(17.0f, 9.0f));
  float field = 0.70f * n0 + 0.30f * n1;
  float cov = smoothstepf(CLOUD_COVERAGE, CLOUD_COVERAGE + CLOUD_SOFTNESS, field);
  return cov * h;
}
// cloud_sun_transmittance_godray, clouds.wgsl:185-215
static inline float cloud_sun_transmittance_godray(const Ctx& C, v3 p, v3 sun_dir) {
  if (sun_dir.y <= 0.01f) return 1.0f;
  float t0 = (CLOUD_BASE_H - p.y) / sun_dir.y, t1 = (CLOUD_TOP_H - p.y) / sun_dir.y;
  float ta = mn(t0, t1), tb = mx(t0, t1);
  if (tb <= 0.0f) return 1.0f;
  float t_start = mx(ta, 0.0f), t_end = tb;
  if (t_end <= t_start) return 1.0f;
  float time_s = C.cam->voxel_params[1];
  float tm = t_start + 0.5f * (t_end - t_start);
  v3 pm = p + sun_dir * tm;
  float dens = cloud_density_godray(pm, time_s);
  float len = (t_end - t_start);
  return vx_exp(-CLOUD_DENSITY * dens * len * CLOUD_SHADOW_ABSORB);
}
// phase_mie / phase_blended, phase.wgsl:5-16
static inline float phase_mie(float costh) {
  float g = PHASE_G, gg = g * g;
  float denom = vx_pow(1.0f + gg - 2.0f * g * costh, 1.5f);
  return INV_4PI * (1.0f - gg) / mx(denom, 1e-3f);
}
static inline float phase_blended(float costh) { return mixf(INV_4PI, phase_mie(costh), PHASE_MIE_W); }
// sky_bg, sky.wgsl:5-25
static inline v3 sky_bg(v3 rd) {
  float t0 = smoothstepf(-0.10f, 0.85f, rd.y);
  v3 col = mix3(V3(0.05f, 0.08f, 0.12f), V3(0.55f, 0.75f, 0.95f), t0);
  float th = 1.0f - smoothstepf(-0.20f, 0.18f, std::fabs(rd.y));
  col = col + (0.12f * th) * V3(0.80f, 0.75f, 0.70f);
  float tz = smoothstepf(0.15f, 1.00f, rd.y);
  col = col + (0.05f * tz) * V3(0.20f, 0.35f, 0.60f);
  col = col * SKY_EXPOSURE;
  return col;
}
// sky_sun, sky.wgsl:27-41
static inline v3 sky_sun(v3 rd) {
  float mu = clampf(dot3(rd, SUN_DIR), -1.0f, 1.0f);
  const float R = 0.009f, S = 0.004f;
  float mu_inner = vx_cos(R), mu_outer = vx_cos(R + S);
  float disc = smoothstepf(mu_outer, mu_inner, mu);
  float glow = smoothstepf(0.70f, 0.98f, mu);
  float halo = 0.03f * vx_exp(-2500.0f * mx(0.0f, 1.0f - mu));
  return SUN_COLOR * SUN_INTENSITY * (disc + halo + 0.04f * glow);
}
// sky_color, sky.wgsl:43-137
static inline v3 sky_color(const Ctx& C, v3 rd) {
  v3 bg = sky_bg(rd);
  v3 sun = sky_sun(rd);
  float T_view = 1.0f;
  v3 acc = V3s(0.0f);
  if (rd.y > 0.01f) {
    v3 ro = cam_pos(C);
    float time_s = C.cam->voxel_params[1];
    float t0 = (CLOUD_BASE_H - ro.y) / rd.y, t1 = (CLOUD_TOP_H - ro.y) / rd.y;
    float t_enter = mn(t0, t1), t_exit = mx(t0, t1);
    t_enter = mx(t_enter, 0.0f);
    if (t_exit > t_enter) {
      float horizon2 = clampf((rd.y - CLOUD_HORIZON_Y0) / CLOUD_HORIZON_Y1, 0.0f, 1.0f);
      float dt = (t_exit - t_enter) / (float)CLOUD_STEPS_VIEW;
      float phase = phase_blended(dot3(rd, SUN_DIR));
      for (uint32_t i = 0; i < CLOUD_STEPS_VIEW; ++i) {
        float ti = t_enter + ((float)i + 0.5f) * dt;
        v3 p = ro + rd * ti;
        float dens = cloud_density(p, time_s);
        dens *= horizon2;
        if (dens > 1e-4f) {
          float Tl = cloud_light_transmittance(p, time_s);
          float toward_sun = clampf(dot3(rd, SUN_DIR), 0.0f, 1.0f);
          float silver = vx_pow(toward_sun, CLOUD_SILVER_POW) * CLOUD_SILVER_STR;
          v3 cloud_col = mix3(CLOUD_BASE_COL, V3s(1.0f), silver);
          float tau = CLOUD_DENSITY * dens * dt;
          float a = 1.0f - vx_exp(-tau);
          v3 scatter = cloud_col * (SUN_COLOR * SUN_INTENSITY) * (phase * Tl);
          acc = acc + T_view * a * scatter;
          T_view *= (1.0f - a);
          if (T_view < 0.02f) break;
        }
      }
    }
  }
  v3 col = bg * T_view + acc;
  float T_sun = 1.0f;
  if (rd.y > 0.01f) {
    T_sun = vx_pow(clampf(T_view, 0.0f, 1.0f), CLOUD_SUN_DISC_DIM_POW);
    T_sun = mx(T_sun, CLOUD_SUN_DISC_DIM_FLOOR);
  }
  col = col + sun * T_sun;
  return col;
}

// ===========================================================================
// ray/fog.wgsl
// ===========================================================================
// fog_color_from_sky, fog.wgsl:5-12
static inline v3 fog_color_from_sky(v3 rd, v3 sky) {
  float up = clampf(rd.y * 0.5f + 0.5f, 0.0f, 1.0f);
  v3 sky_clamped = mn3(sky, V3s(0.45f));
  return mix3(FOG_COLOR_GROUND, sky_clamped, FOG_COLOR_SKY_BLEND * up);
}
// fog_inscatter, fog.wgsl:15-19
static inline v3 fog_inscatter(v3 rd, v3 fogc) {
  float mu = clampf(dot3(rd, SUN_DIR), 0.0f, 1.0f);
  float sun_scatter = vx_pow(mu, 8.0f);
  return fogc + 0.35f * sun_scatter * SUN_COLOR;
}
// fog_density_primary / fog_density_godray, fog.wgsl:21-29
static inline float fog_density_primary(const Ctx& C) { return mx(C.cam->voxel_params[3] * FOG_PRIMARY_SCALE, 0.0f); }
static inline float fog_density_godray(const Ctx& C) { return mx(C.cam->voxel_params[3] * FOG_GODRAY_SCALE, 0.0f); }
// fog_optical_depth_with_base, fog.wgsl:31-45
static inline float fog_optical_depth_with_base(float base, v3 ro, v3 rd, float t) {
  if (base <= 0.0f) return 0.0f;
  float k = FOG_HEIGHT_FALLOFF, y0 = ro.y, dy = rd.y;
  if (std::fabs(dy) < 1e-4f) return base * vx_exp(-k * y0) * t;
  float a = vx_exp(-k * y0);
  float b = vx_exp(-k * (y0 + dy * t));
  return base * (a - b) / (k * dy);
}
// apply_fog, fog.wgsl:57-71 (+ fog_transmittance_primary 47-50)
static inline v3 apply_fog(const Ctx& C, v3 surface, v3 ro, v3 rd, float t_scene, v3 sky) {
  float od = mx(fog_optical_depth_with_base(fog_density_primary(C), ro, rd, t_scene), 0.0f);
  float T = vx_exp(-od);
  v3 fogc = fog_color_from_sky(rd, sky);
  v3 ins = fog_inscatter(rd, fogc);
  return surface * T + ins * (1.0f - T);
}

// ===========================================================================
// ray/aabb.wgsl
// ===========================================================================
struct CubeSlab { v3 tminv, tmaxv; float t_enter, t_exit; };
// cube_slab_inv, aabb.wgsl:16-84
static inline CubeSlab cube_slab_inv(v3 ro, v3 inv, v3 bmin, float size) {
  v3 bmax = bmin + V3s(size);
  v3 t0 = (bmin - ro) * inv;
  v3 t1 = (bmax - ro) * inv;
  const float INV_PARALLEL = 1e20f;
  const CubeSlab miss = {V3s(BIG_F32), V3s(-BIG_F32), BIG_F32, -BIG_F32};
  if (std::fabs(inv.x) > INV_PARALLEL) {
    if (ro.x < bmin.x || ro.x >= bmax.x) return miss;
    t0.x = -BIG_F32;
    t1.x = BIG_F32;
  }
  if (std::fabs(inv.y) > INV_PARALLEL) {
    if (ro.y < bmin.y || ro.y >= bmax.y) return miss;
    t0.y = -BIG_F32;
    t1.y = BIG_F32;
  }
  if (std::fabs(inv.z) > INV_PARALLEL) {
    if (ro.z < bmin.z || ro.z >= bmax.z) return miss;
    t0.z = -BIG_F32;
    t1.z = BIG_F32;
  }
  CubeSlab s;
  s.tminv = mn3(t0, t1);
  s.tmaxv = mx3(t0, t1);
  s.t_enter = mx(s.tminv.x, mx(s.tminv.y, s.tminv.z));
  s.t_exit = mn(s.tmaxv.x, mn(s.tmaxv.y, s.tmaxv.z));
  return s;
}
struct BoxHit { bool hit; float t, t_exit; v3 n; };
// cube_hit_normal_from_slab, aabb.wgsl:95-141
static inline BoxHit cube_hit_normal_from_slab(v3 rd, const CubeSlab& slab, float t_min, float t_max) {
  float t0 = mx(slab.t_enter, t_min);
  if (slab.t_exit < t0 || t0 > t_max) return BoxHit{false, BIG_F32, slab.t_exit, V3s(0.0f)};
  float te = slab.t_enter;
  float eps = 1e-5f * mx(1.0f, std::fabs(te));
  bool bx = std::fabs(slab.tminv.x - te) <= eps;
  bool by = std::fabs(slab.tminv.y - te) <= eps;
  bool bz = std::fabs(slab.tminv.z - te) <= eps;
  uint32_t axis = 0;
  if (bx || by || bz) {
    float best = -1.0f;
    if (bx) { best = std::fabs(rd.x); axis = 0; }
    if (by && std::fabs(rd.y) > best) { best = std::fabs(rd.y); axis = 1; }
    if (bz && std::fabs(rd.z) > best) { best = std::fabs(rd.z); axis = 2; }
  } else {
    float bestv = slab.tminv.x;
    axis = 0;
    if (slab.tminv.y > bestv) { bestv = slab.tminv.y; axis = 1; }
    if (slab.tminv.z > bestv) { bestv = slab.tminv.z; axis = 2; }
  }
  v3 n = V3s(0.0f);
  if (axis == 0) n = V3(rd.x > 0.0f ? -1.0f : 1.0f, 0.0f, 0.0f);
  if (axis == 1) n = V3(0.0f, rd.y > 0.0f ? -1.0f : 1.0f, 0.0f);
  if (axis == 2) n = V3(0.0f, 0.0f, rd.z > 0.0f ? -1.0f : 1.0f);
  return BoxHit{true, t0, slab.t_exit, n};
}
// aabb_hit_normal_inv, aabb.wgsl:144-155
static inline BoxHit aabb_hit_normal_inv(v3 ro, v3 rd, v3 inv, v3 bmin, float size, float t_min, float t_max) {
  CubeSlab s = cube_slab_inv(ro, inv, bmin, size);
  return cube_hit_normal_from_slab(rd, s, t_min, t_max);
}

// ===========================================================================
// ray/wind.wgsl, ray/svo_query.wgsl, ray/leaves.wgsl
// ===========================================================================
// hash1, wind.wgsl:5-8
static inline float hash1(v3 p) {
  float h = dot3(p, V3(127.1f, 311.7f, 74.7f));
  return fractf(vx_sin(h) * 43758.5453f);
}
// wind_field, wind.wgsl:10-36
static inline v3 wind_field(v3 pos_m, float t) {
  v3 cell = floor3(pos_m * WIND_CELL_FREQ);
  float ph0 = hash1(cell);
  float ph1 = hash1(cell + WIND_PHASE_OFF_1);
  v2 dir = normalize2(WIND_DIR_XZ);
  float h = clampf((pos_m.y - WIND_RAMP_Y0) / mx(WIND_RAMP_Y1 - WIND_RAMP_Y0, 1e-3f), 0.0f, 1.0f);
  float gust = vx_sin(t * WIND_GUST_TIME_FREQ + dot2(V2(pos_m.x, pos_m.z), WIND_GUST_XZ_FREQ) + ph0 * TAU);
  float flutter = vx_sin(t * WIND_FLUTTER_TIME_FREQ + dot2(V2(pos_m.x, pos_m.z), WIND_FLUTTER_XZ_FREQ) + ph1 * TAU);
  v2 xz = dir * (WIND_GUST_WEIGHT * gust + WIND_FLUTTER_WEIGHT * flutter) * h;
  float y = WIND_VERTICAL_SCALE * flutter * h;
  return V3(xz.x, y, xz.y);
}
// clamp_len, wind.wgsl:38-42
static inline v3 clamp_len(v3 v, float max_len) {
  float l2 = dot3(v, v);
  if (l2 <= max_len * max_len) return v;
  return v * (max_len / std::sqrt(l2));
}
struct LeafQuery { v3 bmin; float size; uint32_t mat; };
// query_leaf_at, svo_query.wgsl:11-108
static inline LeafQuery query_leaf_at(const Ctx& C, v3 p_in, v3 root_bmin, float root_size, uint32_t node_base, uint32_t macro_base) {
  uint32_t idx = node_base;
  v3 bmin = root_bmin;
  float size = root_size;
  if (macro_base != INVALID_U32) {
    float cell = root_size / (float)MACRO_DIM;
    v3 lp = p_in - root_bmin;
    if (lp.x < 0.0f || lp.y < 0.0f || lp.z < 0.0f || lp.x >= root_size || lp.y >= root_size || lp.z >= root_size)
      return LeafQuery{root_bmin, root_size, MAT_AIR};
    int32_t mxi = f2i(std::floor(lp.x / cell)), myi = f2i(std::floor(lp.y / cell)), mzi = f2i(std::floor(lp.z / cell));
    if (mxi < 0 || myi < 0 || mzi < 0 || mxi >= (int32_t)MACRO_DIM || myi >= (int32_t)MACRO_DIM || mzi >= (int32_t)MACRO_DIM)
      return LeafQuery{root_bmin, root_size, MAT_AIR};
    uint32_t bit = macro_bit_index((uint32_t)mxi, (uint32_t)myi, (uint32_t)mzi);
    if (!macro_test(C, macro_base, bit)) return LeafQuery{root_bmin, root_size, MAT_AIR};
  }
  v3 p = p_in;
  for (uint32_t d = 0; d < 32u; ++d) {
    vx_node n = ld_node(C, idx);
    if (n.child_base == LEAF_U32) return LeafQuery{bmin, size, n.material};
    uint32_t max_d = chunk_max_depth(C);
    if (d >= max_d) return LeafQuery{bmin, size, MAT_AIR};
    float half = size * 0.5f;
    v3 mid = bmin + V3s(half);
    float e = 1e-6f * size;
    uint32_t hx = p.x > mid.x + e ? 1u : 0u, hy = p.y > mid.y + e ? 1u : 0u, hz = p.z > mid.z + e ? 1u : 0u;
    uint32_t ci = hx | (hy << 1u) | (hz << 2u);
    v3 child_bmin = bmin + V3(hx ? half : 0.0f, hy ? half : 0.0f, hz ? half : 0.0f);
    uint32_t bit = 1u << ci;
    if ((n.child_mask & bit) == 0u) return LeafQuery{child_bmin, half, MAT_AIR};
    uint32_t rank = child_rank(n.child_mask, ci);
    idx = node_base + (n.child_base + rank);
    bmin = child_bmin;
    size = half;
  }
  return LeafQuery{bmin, size, MAT_AIR};
}
// query_leaf_world, svo_query.wgsl:115-136
static inline LeafQuery query_leaf_world(const Ctx& C, v3 p) {
  float vs = vp_x(C);
  float csm = (float)C.cam->chunk_size * vs;
  i3 c = chunk_coord_from_pos(p, csm);
  uint32_t slot = grid_lookup_slot(C, c.x, c.y, c.z);
  if (slot == INVALID_U32 || slot >= C.cam->chunk_count) return LeafQuery{p, vs, MAT_AIR};
  vx_chunk_meta ch = ld_chunk(C, slot);
  v3 root_bmin = V3((float)ch.origin[0], (float)ch.origin[1], (float)ch.origin[2]) * vs;
  float root_size = (float)C.cam->chunk_size * vs;
  return query_leaf_at(C, p, root_bmin, root_size, ch.node_base, ch.macro_base);
}
// leaf_cube_offset, leaves.wgsl:5-17
static inline v3 leaf_cube_offset(const Ctx& C, v3 bmin, float size, float time_s, float strength) {
  v3 center = bmin + V3s(0.5f * size);
  float d = length3(center - cam_pos(C));
  float fade = smoothstepf(LEAF_LOD_DISP_START, LEAF_LOD_DISP_END, d);
  float strength_eff = strength * (1.0f - fade);
  v3 w = wind_field(center, time_s) * strength_eff;
  w = V3(w.x, LEAF_VERTICAL_REDUCE * w.y, w.z);
  float amp = LEAF_OFFSET_AMP * size;
  return clamp_len(w * amp, LEAF_OFFSET_MAX_FRAC * size);
}
struct LeafCubeHit { bool hit; float t; v3 n; };
// leaf_displaced_cube_hit, leaves.wgsl:26-43
static inline LeafCubeHit leaf_displaced_cube_hit(const Ctx& C, v3 ro, v3 rd, v3 bmin, float size, float time_s,
                                                  float strength, float t_min, float t_max) {
  v3 off = leaf_cube_offset(C, bmin, size, time_s, strength);
  v3 bmin2 = bmin + off;
  v3 inv = safe_inv3(rd);
  BoxHit bh = aabb_hit_normal_inv(ro, rd, inv, bmin2, size, t_min, t_max);
  return LeafCubeHit{bh.hit, bh.t, bh.n};
}

// ===========================================================================
// ray/grass.wgsl
// ===========================================================================
struct GrassCell { v3 bmin_m, id_vox; };
// pick_grass_cell_in_chunk_with_bias, grass.wgsl:10-44
static inline GrassCell pick_grass_cell_in_chunk_with_bias(v3 hp, v3 bias_dir, v3 root_bmin, i3 org, float vs, int32_t csv) {
  float root_size_m = (float)csv * vs;
  float bias = 0.05f * vs;
  v3 lxz = (hp - root_bmin) - bias_dir * bias;
  lxz.x = clampf(lxz.x, 0.0f, root_size_m - 1e-6f);
  lxz.z = clampf(lxz.z, 0.0f, root_size_m - 1e-6f);
  float y_in = (hp - bias_dir * bias).y - 1e-4f * vs;
  float ly = clampf(y_in - root_bmin.y, 0.0f, root_size_m - 1e-6f);
  int32_t ix = clampi(f2i(std::floor(lxz.x / vs)), 0, csv - 1);
  int32_t iy = clampi(f2i(std::floor(ly / vs)), 0, csv - 1);
  int32_t iz = clampi(f2i(std::floor(lxz.z / vs)), 0, csv - 1);
  GrassCell g;
  g.bmin_m = root_bmin + V3((float)ix, (float)iy, (float)iz) * vs;
  g.id_vox = V3((float)(org.x + ix), (float)(org.y + iy), (float)(org.z + iz));
  return g;
}
// grass_cell_from_world, grass.wgsl:64-84
static inline GrassCell grass_cell_from_world(v3 hp, v3 rd, v3 root_bmin, float vs, int32_t csv) {
  i3 org{f2i(std::floor(root_bmin.x / vs)), f2i(std::floor(root_bmin.y / vs)), f2i(std::floor(root_bmin.z / vs))};
  return pick_grass_cell_in_chunk_with_bias(hp, rd, root_bmin, org, vs, csv);
}
// grass_cell_from_world_global, grass.wgsl:86-104
static inline GrassCell grass_cell_from_world_global(v3 hp, v3 rd, float vs) {
  float bias = 0.05f * vs;
  v3 p = hp - rd * bias;
  float py = p.y - 1e-4f * vs;
  int32_t ix = f2i(std::floor(p.x / vs)), iy = f2i(std::floor(py / vs)), iz = f2i(std::floor(p.z / vs));
  GrassCell g;
  g.bmin_m = V3((float)ix, (float)iy, (float)iz) * vs;
  g.id_vox = V3((float)ix, (float)iy, (float)iz);
  return g;
}
// sdf_box, grass.wgsl:106-111
static inline float sdf_box(v3 p, v3 c, v3 b) {
  v3 q = abs3(p - c) - b;
  float outside = length3(mx3(q, V3s(0.0f)));
  float inside = mn(mx(q.x, mx(q.y, q.z)), 0.0f);
  return outside + inside;
}
// make_orthonormal_basis column 0, grass.wgsl:120-126
static inline v3 orthonormal_basis_x(v3 n) {
  v3 up = std::fabs(n.y) > 0.999f ? V3(1.0f, 0.0f, 0.0f) : V3(0.0f, 1.0f, 0.0f);
  return normalize3(cross3(up, n));
}
// grass_wind_xz, grass.wgsl:185-189
static inline v2 grass_wind_xz(v3 root, float t, float strength) {
  v3 w = wind_field(root, t) * strength;
  return V2(w.x, w.z);
}
// grass_blade_params, grass.wgsl:353-368
static inline v4 grass_blade_params(v3 id, uint32_t i) {
  float fi = (float)i;
  float h0 = hash31(id + V3(fi * 7.3f, 1.1f, 2.9f));
  float h1 = hash31(id + V3(fi * 3.7f, 8.4f, 0.6f));
  float u = fractf(h0 * 1.61803398875f);
  float v = fractf(h0 * 2.41421356237f);
  float p = fractf(h0 * 3.14159265359f);
  float w = 0.35f + 0.65f * (h1 * h1);
  return v4{u, v, p, w};
}
// grass_sdf_lod, grass.wgsl:192-350
static inline float grass_sdf_lod(const Ctx& C, v3 p_m, v3 cell_bmin, v3 id, float time_s, float strength, uint32_t lod) {
  float vs = vp_x(C);
  float top_y = cell_bmin.y + vs;
  float layer_h = GRASS_LAYER_HEIGHT_VOX * vs;
  float y01 = (p_m.y - top_y) / mx(layer_h, 1e-6f);
  if (y01 < 0.0f || y01 > 1.0f) return BIG_F32;
  float over = GRASS_OVERHANG_VOX * vs;
  if (p_m.x < cell_bmin.x - over || p_m.x > cell_bmin.x + vs + over || p_m.z < cell_bmin.z - over ||
      p_m.z > cell_bmin.z + vs + over)
    return BIG_F32;
  float blade_len = layer_h * (0.65f + 0.35f * hash31(id + V3(9.1f, 3.7f, 5.2f)));
  uint32_t blade_count = GRASS_BLADE_COUNT;
  uint32_t segs = f2u(mx(3.0f, std::floor(GRASS_VOXEL_SEGS)));
  if (lod == 1u) {
    blade_count = blade_count < GRASS_BLADE_COUNT_MID ? blade_count : GRASS_BLADE_COUNT_MID;
    segs = segs < GRASS_SEGS_MID ? segs : GRASS_SEGS_MID;
  } else if (lod == 2u) {
    blade_count = blade_count < GRASS_BLADE_COUNT_FAR ? blade_count : GRASS_BLADE_COUNT_FAR;
    segs = segs < GRASS_SEGS_FAR ? segs : GRASS_SEGS_FAR;
  }
  float inv_segs = 1.0f / mx((float)segs, 1.0f);
  float dmin = BIG_F32;
  const uint32_t CLUMPS = 5;
  float clump_r = 0.18f * vs;
  float base_r = (GRASS_VOXEL_THICKNESS_VOX * vs) * 0.85f;
  for (uint32_t i = 0; i < blade_count; ++i) {
    v4 uvpw = grass_blade_params(id, i);
    float u0 = uvpw.x, v0 = uvpw.y, ph = uvpw.z, wR = uvpw.w;
    uint32_t ci = i % CLUMPS;
    float cfi = (float)ci;
    float cu = hash31(id + V3(13.7f + cfi, 2.1f, 9.9f));
    float cv = hash31(id + V3(4.2f + cfi, 7.3f, 1.4f));
    const float inset = 0.10f;
    float cux = mixf(inset, 1.0f - inset, cu);
    float cvz = mixf(inset, 1.0f - inset, cv);
    v2 clump_center = V2(cell_bmin.x + cux * vs, cell_bmin.z + cvz * vs);
    float ang = TAU * u0;
    float rr = clump_r * (0.15f + 0.85f * v0);
    v2 root_xz = clump_center + V2(vx_cos(ang), vx_sin(ang)) * rr;
    v3 root = V3(root_xz.x, top_y, root_xz.y);
    v2 w_xz = 0.35f * grass_wind_xz(root + V3(0.0f, ph, 0.0f), time_s, strength);
    float lean_ang = TAU * fractf(u0 * 1.73f + v0 * 2.11f);
    v2 lean_dir = V2(vx_cos(lean_ang), vx_sin(lean_ang));
    float lean_amt = 0.10f + 0.25f * (v0 * v0);
    float r0 = base_r * mixf(0.85f, 1.35f, wR);
    float roll = TAU * (0.15f * ph + 0.10f * wR);
    v2 roll_dir = V2(vx_cos(roll), vx_sin(roll));
    v3 side_hint = normalize3(V3(lean_dir.x * roll_dir.x - lean_dir.y * roll_dir.y, 0.12f,
                                 lean_dir.x * roll_dir.y + lean_dir.y * roll_dir.x));
    for (uint32_t s = 0; s < segs; ++s) {
      float t01a = ((float)s) * inv_segs;
      float t01b = ((float)s + 1.0f) * inv_segs;
      float ya = t01a * blade_len, yb = t01b * blade_len;
      float seg_h = yb - ya;
      float bend_a = (blade_len * t01a) * (0.55f + 0.45f * t01a);
      float bend_b = (blade_len * t01b) * (0.55f + 0.45f * t01b);
      v2 bva = w_xz * t01a + lean_dir * lean_amt;
      v2 bvb = w_xz * t01b + lean_dir * lean_amt;
      v2 offa = bva * bend_a, offb = bvb * bend_b;
      v3 pa = root + V3(offa.x, ya, offa.y);
      v3 pb = root + V3(offb.x, yb, offb.y);
      v3 T = normalize3(pb - pa);
      v3 S = side_hint - T * dot3(side_hint, T);
      float s2 = dot3(S, S);
      if (s2 < 1e-6f) S = orthonormal_basis_x(T);
      else S = S * inverse_sqrt(s2);
      v3 N = cross3(T, S);
      v3 mid = 0.5f * (pa + pb);
      v3 d = p_m - mid;
      v3 local = V3(dot3(d, S), dot3(d, T), dot3(d, N));
      float tmid = 0.5f * (t01a + t01b);
      float taper_shape = vx_pow(clampf(1.0f - tmid, 0.0f, 1.0f), 3.0f);
      float taper = mixf(GRASS_VOXEL_TAPER, 1.0f, taper_shape);
      float half_w = r0 * taper * 0.7f;
      float half_t = half_w;
      float dbox = sdf_box(local, V3s(0.0f), V3(half_w, 0.5f * seg_h, half_t));
      dmin = mn(dmin, dbox);
    }
  }
  return dmin;
}
// grass_sdf_normal_lod, grass.wgsl:370-408
static inline v3 grass_sdf_normal_lod(const Ctx& C, v3 p, v3 cb, v3 id, float ts, float st, uint32_t lod) {
  if (lod != 0u) {
    float vs = vp_x(C);
    float top_y = cb.y + vs;
    v3 root = V3(cb.x + 0.5f * vs, top_y, cb.z + 0.5f * vs);
    v2 w = grass_wind_xz(root, ts, st);
    const float tilt = 0.35f;
    return normalize3(V3(-w.x * tilt, 1.0f, -w.y * tilt));
  }
  float e = 0.01f * vp_x(C);
  float dx = grass_sdf_lod(C, p + V3(e, 0.0f, 0.0f), cb, id, ts, st, lod) - grass_sdf_lod(C, p - V3(e, 0.0f, 0.0f), cb, id, ts, st, lod);
  float dy = grass_sdf_lod(C, p + V3(0.0f, e, 0.0f), cb, id, ts, st, lod) - grass_sdf_lod(C, p - V3(0.0f, e, 0.0f), cb, id, ts, st, lod);
  float dz = grass_sdf_lod(C, p + V3(0.0f, 0.0f, e), cb, id, ts, st, lod) - grass_sdf_lod(C, p - V3(0.0f, 0.0f, e), cb, id, ts, st, lod);
  return normalize3(V3(dx, dy, dz));
}
// grass_self_shadow, grass.wgsl:410-447
static inline float grass_self_shadow(const Ctx& C, v3 hp, v3 cb, v3 id, float ts, float st, uint32_t lod) {
  float vs = vp_x(C);
  v3 p = hp + SUN_DIR * (0.02f * vs);
  float max_dist = (0.55f + 0.35f * (float)lod) * vs;
  uint32_t steps = lod != 0u ? 4u : 6u;
  float dt = max_dist / (float)steps;
  float occ = 0.0f;
  for (uint32_t i = 0; i < steps; ++i) {
    float d = grass_sdf_lod(C, p, cb, id, ts, st, lod);
    float x = clampf(1.0f - d / (0.05f * vs), 0.0f, 1.0f);
    occ += x;
    p = p + SUN_DIR * dt;
  }
  return vx_exp(-0.35f * occ);
}
struct GrassHit { bool hit; float t; v3 n; };
// grass_lod_from_t, grass.wgsl:540-545
static inline uint32_t grass_lod_from_t(float t) {
  if (t >= GRASS_LOD_FAR_START) return 2u;
  if (t >= GRASS_LOD_MID_START) return 1u;
  return 0u;
}
// grass_layer_trace_lod, grass.wgsl:456-491
static inline GrassHit grass_layer_trace_lod(const Ctx& C, v3 ro, v3 rd, float t_start, float t_end, v3 cb, v3 id,
                                             float ts, float st, uint32_t lod) {
  float vs = vp_x(C);
  float t = t_start;
  uint32_t steps = GRASS_TRACE_STEPS;
  if (lod == 1u) steps = GRASS_TRACE_STEPS_MID;
  if (lod == 2u) steps = GRASS_TRACE_STEPS_FAR;
  for (uint32_t i = 0; i < steps; ++i) {
    if (t > t_end) break;
    v3 p = ro + rd * t;
    float d = grass_sdf_lod(C, p, cb, id, ts, st, lod);
    float hit_eps = GRASS_HIT_EPS_VOX * vs;
    if (d < hit_eps) {
      v3 n = grass_sdf_normal_lod(C, p, cb, id, ts, st, lod);
      return GrassHit{true, t, n};
    }
    float step_min = GRASS_STEP_MIN_VOX * vs;
    t += mx(d, step_min);
  }
  return GrassHit{false, BIG_F32, V3s(0.0f)};
}
// try_grass_slab_hit, grass.wgsl:494-538
static inline GrassHit try_grass_slab_hit(const Ctx& C, v3 ro, v3 rd, float t_min, float t_max, v3 cb, v3 id, float vs,
                                          float ts, float st) {
  float layer_h = GRASS_LAYER_HEIGHT_VOX * vs;
  float over = GRASS_OVERHANG_VOX * vs;
  v3 slab_bmin = V3(cb.x - over, cb.y + vs, cb.z - over);
  v3 slab_bmax = V3(cb.x + vs + over, cb.y + vs + layer_h, cb.z + vs + over);
  v2 rt = intersect_aabb(ro, rd, slab_bmin, slab_bmax);
  float t0 = mx(rt.x, t_min);
  float t1 = mn(rt.y, t_max);
  v3 vox_bmax = cb + V3s(vs);
  v2 rv = intersect_aabb(ro, rd, cb, vox_bmax);
  float clip_eps = 0.01f * vs;
  if (rv.y > rv.x) {
    float tev = rv.x;
    if (tev > t0) t1 = mn(t1, tev - clip_eps);
  }
  if (t1 <= t0) return GrassHit{false, BIG_F32, V3s(0.0f)};
  uint32_t lod = grass_lod_from_t(t0);
  return grass_layer_trace_lod(C, ro, rd, t0, t1, cb, id, ts, st, lod);
}

}  // namespace vxo
