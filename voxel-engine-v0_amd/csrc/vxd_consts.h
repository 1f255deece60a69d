// vxd_consts.h — tuning constants of the passes, values as in the reference
// (src/shaders/common.wgsl:16-332 and the function-local constants of
// ray/sky.wgsl, ray/composite.wgsl, ray/grass.wgsl).
#pragma once
#include "vxd_math.h"

namespace vxd {

constexpr uint32_t LEAF_U32 = 0xFFFFFFFFu;
constexpr uint32_t MAT_AIR = 0, MAT_GRASS = 1, MAT_DIRT = 2, MAT_STONE = 3, MAT_WOOD = 4, MAT_LEAF = 5, MAT_LIGHT = 6;
#define VXD_SUN_DIR F3(0.61237244f, 0.5f, 0.61237244f)
#define VXD_SUN_COLOR F3(1.0f, 0.94f, 0.72f)
#define VXD_FOG_COLOR_GROUND F3(0.40f, 0.42f, 0.45f)
constexpr float SUN_INTENSITY = 5.0f;
constexpr float SHADOW_BIAS = 2e-4f;
constexpr uint32_t VSM_STEPS = 16;
constexpr float LEAF_LIGHT_TRANSMIT = 0.50f, GRASS_LIGHT_TRANSMIT = 0.70f, MIN_TRANS = 0.03f;
constexpr float GODRAY_ENERGY_BOOST = 8.0f, GODRAY_KNEE_INTEGRATE = 0.35f, GODRAY_COMPOSITE_SCALE = 6.5f;
constexpr float GODRAY_FADE_NEAR = 40.0f, GODRAY_FADE_FAR = 80.0f, GODRAY_KNEE_COMPOSITE = 0.25f;
constexpr float FOG_HEIGHT_FALLOFF = 0.18f, FOG_MAX_DIST = 120.0f, FOG_PRIMARY_SCALE = 0.02f, FOG_GODRAY_SCALE = 1.0f;
constexpr float FOG_COLOR_SKY_BLEND = 0.10f;
constexpr float GODRAY_MAX_DIST = 100.0f, GODRAY_SCATTER_HEIGHT_FALLOFF = 0.04f, GODRAY_SCATTER_MIN_FRAC = 0.35f;
constexpr float GODRAY_BLACK_LEVEL = 0.004f, GODRAY_EDGE0 = 0.004f, GODRAY_EDGE1 = 0.035f;
constexpr float GODRAY_BASE_HAZE = 0.02f, GODRAY_HAZE_NEAR_FADE = 18.0f, GODRAY_TV_CUTOFF = 0.02f;
constexpr uint32_t GODRAY_STEPS_FAST = 24, GODRAY_MIN_STEPS = 8, GODRAY_STEP_Q = 4;
constexpr float GODRAY_STEPS_PER_METER = 1.25f, GODRAY_EDGE_ENERGY_BOOST = 2.5f;
constexpr float INV_4PI = 0.0795774715f, PHASE_G = 0.15f, PHASE_MIE_W = 0.20f;
constexpr float CLOUD_BASE_H = 170.0f, CLOUD_TOP_H = 250.0f, CLOUD_UV_SCALE = 0.0016f;
constexpr float CLOUD_WIND_X = 0.020f, CLOUD_WIND_Y = 0.012f;
constexpr float CLOUD_COVERAGE = 0.40f, CLOUD_SOFTNESS = 0.08f, CLOUD_DENSITY = 0.058f;
constexpr float CLOUD_PUFF_POW = 1.8f, CLOUD_DETAIL_W = 0.35f;
constexpr uint32_t CLOUD_STEPS_VIEW = 8, CLOUD_STEPS_LIGHT = 4;
constexpr float CLOUD_HORIZON_Y0 = 0.02f, CLOUD_HORIZON_Y1 = 0.25f;
constexpr float CLOUD_SILVER_POW = 8.0f, CLOUD_SILVER_STR = 0.6f;
constexpr float CLOUD_SHADOW_ABSORB = 6.0f;
constexpr float CLOUD_SUN_DISC_DIM_POW = 8.0f, CLOUD_SUN_DISC_DIM_FLOOR = 0.0f;
constexpr float CLOUD_GODRAY_W = 0.50f, SKY_EXPOSURE = 0.40f;
constexpr float WIND_CELL_FREQ = 2.5f;
constexpr float WIND_RAMP_Y0 = 2.0f, WIND_RAMP_Y1 = 14.0f;
constexpr float WIND_GUST_TIME_FREQ = 0.9f, WIND_FLUTTER_TIME_FREQ = 4.2f;
constexpr float WIND_GUST_WEIGHT = 0.75f, WIND_FLUTTER_WEIGHT = 0.25f;
constexpr float WIND_VERTICAL_SCALE = 0.25f, LEAF_VERTICAL_REDUCE = 0.15f;
constexpr float LEAF_OFFSET_AMP = 0.75f, LEAF_OFFSET_MAX_FRAC = 0.75f;
constexpr float LEAF_LOD_DISP_START = 25.0f, LEAF_LOD_DISP_END = 70.0f;
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

// Constant divisors whose quotients the passes compute with a two-term
// reciprocal (divc below) instead of the full division sequence.  For every
// entry divc equals the correctly rounded x / d for every binary32 x: checked
// exhaustively over all non-negative bit patterns (divc is odd in x), so the
// quotients stay bit-identical to a division.  Entries equal the divisor each
// call site's expression yields.
constexpr float FAST_DIVISORS[] = {
    GODRAY_EDGE1 - GODRAY_EDGE0,                         // godray shaft edge smoothstep
    0.80f - 0.35f,                                       // godray sun gate smoothstep
    GODRAY_HAZE_NEAR_FADE,                               // godray haze ramp
    CLOUD_TOP_H - CLOUD_BASE_H,                          // cloud height fraction
    1.00f - 0.78f,                                       // cloud height upper smoothstep
    (CLOUD_COVERAGE + CLOUD_SOFTNESS) - CLOUD_COVERAGE,  // cloud coverage smoothstep
};
enum : uint32_t { DIV_GODRAY_EDGE, DIV_GODRAY_GATE, DIV_HAZE_FADE, DIV_CLOUD_H, DIV_CLOUD_HI, DIV_COVERAGE,
                  DIV_COUNT };
static_assert(sizeof(FAST_DIVISORS) / sizeof(float) == DIV_COUNT, "one divisor per name");

// x / FAST_DIVISORS[K] as fma(x, zh, x * zl), where zh + zl is 1/d to about
// twice binary32 precision: zh = 1/d rounded, zl = (1/d - zh) rounded
// (Brisebarre, Muller, Raina, "Accelerating correctly rounded floating-point
// division when the divisor is known in advance", IEEE TC 53(8), 2004).  For
// |x| below 2^-100 x * zl loses bits (and the method has no guarantee), so
// those inputs, zero, infinities and NaN take the division itself.  Some
// divisors have a few failing x even in range (0.12 has 225); such divisors are
// not in the table.
template <uint32_t K>
VXD float divc(float x) {
  constexpr float d = FAST_DIVISORS[K];
  constexpr float zh = 1.0f / d;
  constexpr float zl = (float)(1.0 / (double)d - (double)zh);
  const float ax = fabsf(x);
  if (__builtin_expect(ax >= 0x1p-100f && ax <= 0x1p100f, 1)) return fmaf(x, zh, x * zl);
  return x / d;
}
// smoothstepf(lo, lo + FAST_DIVISORS[K], x) for a call site whose hi - lo is entry K.
template <uint32_t K>
VXD float smoothstepc(float lo, float x) {
  float t = clampf(divc<K>(x - lo), 0.0f, 1.0f);
  return t * t * (3.0f - 2.0f * t);
}

}  // namespace vxd
