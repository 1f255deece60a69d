// vx_terrain.h — ground height of the seeded synthetic world and the clipmap
// texel sample, one definition for the host producers (g++) and the device
// clipmap fill kernel (hipcc).
//
// Both compilers see the same IEEE operations in the same order: binary64
// value noise (add, multiply, correctly rounded divide, floor) and binary32
// texel centres (multiply, add, correctly rounded divide by the voxel size),
// with contraction off in both builds, so a height is bit-identical wherever
// it is computed.
#pragma once

#include <math.h>
#include <stdint.h>

#include "voxhip.h"

#if defined(__HIPCC__)
#define VX_HD __host__ __device__ inline
#else
#define VX_HD inline
#endif

namespace vxt {

// Murmur3 32-bit finaliser.
VX_HD uint32_t mix32(uint32_t h) {
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  h ^= h >> 16;
  return h;
}

VX_HD uint32_t lattice_hash(int32_t x, int32_t z, uint32_t salt) {
  return mix32((uint32_t)x * 0x9E3779B1u ^ mix32((uint32_t)z * 0x85EBCA77u ^ salt));
}

VX_HD double lattice_value(int32_t x, int32_t z, uint32_t salt) {
  return (double)(lattice_hash(x, z, salt) >> 8) * (1.0 / 16777216.0);
}

// Smoothstep-interpolated lattice noise in [0, 1).
VX_HD double value_noise_2d(double x, double z, uint32_t salt) {
  const double fx = floor(x), fz = floor(z);
  const int32_t ix = (int32_t)fx, iz = (int32_t)fz;
  const double tx = x - fx, tz = z - fz;
  const double ux = tx * tx * (3.0 - 2.0 * tx), uz = tz * tz * (3.0 - 2.0 * tz);
  const double a = lattice_value(ix, iz, salt), b = lattice_value(ix + 1, iz, salt);
  const double c = lattice_value(ix, iz + 1, salt), d = lattice_value(ix + 1, iz + 1, salt);
  const double x0 = a + (b - a) * ux, x1 = c + (d - c) * ux;
  return x0 + (x1 - x0) * uz;
}

// Surface voxel y of column (wx, wz): two octaves, wavelengths l0 and l0/2.
VX_HD int32_t ground_height(const vx_world_params& p, int32_t wx, int32_t wz) {
  const double l0 = (double)(p.wavelength0 > 0 ? p.wavelength0 : 64);
  const double o0 = value_noise_2d((double)wx / l0, (double)wz / l0, p.seed * 0x9E3779B9u + 0x1234u);
  const double o1 = value_noise_2d((double)wx / (0.5 * l0), (double)wz / (0.5 * l0), p.seed * 0x9E3779B9u + 0xBEEFu);
  const double n = (o0 + 0.5 * o1) / 1.5;  // [0,1]
  return p.base_y + (int32_t)floor((double)p.amplitude * (2.0 * n - 1.0) + 0.5);
}

// float -> i32 with saturation and NaN -> 0 (the semantics of a float-to-int
// `as` cast in clipmap.rs:86-98).
VX_HD int32_t sat_i32(float v) {
  if (!(v == v)) return 0;
  if (v >= 2147483648.0f) return 2147483647;
  if (v < -2147483648.0f) return (int32_t)0x80000000u;
  return (int32_t)v;
}

constexpr float kVoxelM = 0.1f;  // app/config.rs:10

// Height in metres of logical texel (lx, lz) of a level with origin (ox, oz)
// and cell size `cell`: the ground height of the voxel column under the texel
// centre, times the voxel size (clipmap.rs:93-122).
VX_HD float clip_texel_height(const vx_world_params& p, float ox, float oz, float cell, uint32_t lx, uint32_t lz) {
  const float wx = ox + ((float)lx + 0.5f) * cell;
  const float wz = oz + ((float)lz + 0.5f) * cell;
  const int32_t vx = sat_i32(floorf(wx / kVoxelM));
  const int32_t vz = sat_i32(floorf(wz / kVoxelM));
  return (float)ground_height(p, vx, vz) * kVoxelM;
}

}  // namespace vxt
