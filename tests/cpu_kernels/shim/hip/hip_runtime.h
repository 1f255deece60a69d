// Host stand-in for the few HIP names the device pass code uses, so that
// voxel-engine-v0_amd/csrc/vxd_passes.inc compiles as ordinary C++ for a CPU
// test build.  Test infrastructure only: the product is built by hipcc.
#pragma once
#include <math.h>
#include <stdint.h>
#include <string.h>

#define __device__
#define __host__
#define __forceinline__ inline
#define __restrict__ __restrict

struct float4 { float x, y, z, w; };
struct uint4 { uint32_t x, y, z, w; };
inline float4 make_float4(float x, float y, float z, float w) { return float4{x, y, z, w}; }
inline uint4 make_uint4(uint32_t x, uint32_t y, uint32_t z, uint32_t w) { return uint4{x, y, z, w}; }
inline uint32_t __float_as_uint(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
inline float __uint_as_float(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
inline int __clz(uint32_t x) { return x ? __builtin_clz(x) : 32; }
// The device's approximate square root (v_sqrt_f32); the pass code uses it only for
// conservative culling bounds, so the correctly rounded one stands in here.
#define __builtin_amdgcn_sqrtf(x) sqrtf(x)
inline int __popc(uint32_t x) { return __builtin_popcount(x); }
inline uint32_t min(uint32_t a, uint32_t b) { return a < b ? a : b; }
inline uint32_t max(uint32_t a, uint32_t b) { return a > b ? a : b; }
inline int32_t min(int32_t a, int32_t b) { return a < b ? a : b; }
inline int32_t max(int32_t a, int32_t b) { return a > b ? a : b; }
// LDS atomics: each host task owns its workgroup's LDS, so a plain add is exact.
inline uint32_t atomicAdd(uint32_t* p, uint32_t v) { uint32_t o = *p; *p = o + v; return o; }
// Wavefront primitives: each host lane runs as a wavefront of one lane (the
// lane id is the one the host loop set), so a cross-lane read returns the
// lane's own value and a ballot holds its own bit.
extern thread_local uint32_t vxd_cpu_lane;
#define __builtin_amdgcn_read_exec() (1ull << vxd_cpu_lane)
inline uint32_t __lane_id() { return vxd_cpu_lane; }
inline unsigned long long __ballot(int p) { return p ? 1ull << vxd_cpu_lane : 0ull; }
inline int __shfl(int v, int, int) { return v; }
inline float __shfl(float v, int, int) { return v; }
#define __builtin_amdgcn_ds_permute(addr, v) (v)
inline int __popcll(unsigned long long x) { return __builtin_popcountll(x); }
inline void __threadfence_block() {}
