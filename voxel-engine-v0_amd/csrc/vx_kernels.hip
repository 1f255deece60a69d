// vx_kernels.hip — gfx950 kernels of the primary / local TAA / godray /
// composite / composite TAA passes, and their launchers.
//
// Launch mapping (MI355X: 256 CUs in 8 XCDs, 64-lane wavefronts):
//   * primary and godray: one 64-thread workgroup = one wavefront = one 8x8
//     pixel tile (the reference's @workgroup_size(8,8), ray_main.wgsl:163,
//     whose 8x8 tile also scopes the primary candidate chunk list).
//   * Workgroups are dealt to the 8 XCDs round-robin by the hardware; the
//     kernels remap blockIdx so each XCD owns one contiguous band of tiles,
//     keeping the chunks/nodes a band touches in that XCD's L2.
//   * Streaming passes (local TAA, composite, composite TAA) use 256-thread
//     workgroups over 16x16 pixel blocks.
#include <hip/hip_runtime.h>

#include "vx_launch.h"
#include "vxd_consts.h"
#include "vxd_scene.h"

namespace vxd {
namespace plain {
#define VXD_COUNT 0
#include "vxd_passes.inc"
#undef VXD_COUNT
}  // namespace plain
namespace counted {
#define VXD_COUNT 1
#include "vxd_passes.inc"
#undef VXD_COUNT
}  // namespace counted
}  // namespace vxd

namespace {

using namespace vxd;

// Tile order: XCD x (= blockIdx % 8) gets the contiguous tile range
// [x*per + min(x, rem), ...) of per + (x < rem) tiles.
__device__ __forceinline__ uint32_t xcd_tile(uint32_t bid, uint32_t ntiles) {
  const uint32_t per = ntiles >> 3, rem = ntiles & 7u;
  const uint32_t x = bid & 7u, local = bid >> 3;
  return x * per + min(x, rem) + local;
}

template <bool COUNT>
__device__ __forceinline__ void flush_counts(const Counts& c, unsigned long long* out) {
  if (!COUNT) return;
  const uint32_t v[kNumCounts] = {c.grid, c.meta, c.macro, c.node, c.rope, c.colinfo, c.clip, c.tex_rd, c.tex_wr};
#pragma unroll
  for (int i = 0; i < kNumCounts; ++i) {
    unsigned long long s = v[i];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63u) == 0u) atomicAdd(out + i, s);
  }
}

// k_primary is latency bound (dependent node/rope loads): the register budget
// is capped so that 6 waves fit per SIMD (<= 80 VGPRs) instead of 4.
template <bool COUNT>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(6, 8))) void k_primary(
    Scene S, const vx_camera* __restrict__ cam, FrameBuffers fb, uint32_t tiles_x, uint32_t ntiles,
    unsigned long long* __restrict__ counters) {
  __shared__ TileLDS L;
  const uint32_t tile = xcd_tile(blockIdx.x, ntiles);
  const uint32_t wgx = tile % tiles_x, wgy = tile / tiles_x;
  if (threadIdx.x < 12u) L.grass_taper[threadIdx.x] = plain::grass_taper_entry(threadIdx.x);
  if constexpr (COUNT) {
    counted::Ctx C;
    C.S = S;
    C.cam = cam;
    C.grass_taper = L.grass_taper;
    Counts cnt = {};
    C.cnt = &cnt;
    if (threadIdx.x == 0) counted::tile_build(C, L, wgx, wgy);
    __syncthreads();
    const uint32_t gx = wgx * 8u + (threadIdx.x & 7u), gy = wgy * 8u + (threadIdx.x >> 3);
    if (gx < fb.rw && gy < fb.rh)
      counted::primary_pixel(C, L, fb, gx, gy, counted::sky_bg(F3(0.0f, 1.0f, 0.0f)));
    flush_counts<true>(cnt, counters);
  } else {
    plain::Ctx C;
    C.S = S;
    C.cam = cam;
    C.grass_taper = L.grass_taper;
    if (threadIdx.x == 0) plain::tile_build(C, L, wgx, wgy);
    __syncthreads();
    const uint32_t gx = wgx * 8u + (threadIdx.x & 7u), gy = wgy * 8u + (threadIdx.x >> 3);
    if (gx < fb.rw && gy < fb.rh)
      plain::primary_pixel(C, L, fb, gx, gy, plain::sky_bg(F3(0.0f, 1.0f, 0.0f)));
  }
}

__global__ __launch_bounds__(256) void k_local_taa(FrameBuffers fb) {
  const uint32_t x = blockIdx.x * 16u + (threadIdx.x & 15u), y = blockIdx.y * 16u + (threadIdx.x >> 4);
  if (x < fb.rw && y < fb.rh) plain::local_taa_pixel(fb, x, y);
}

template <bool COUNT>
__global__ __launch_bounds__(64) void k_godray(Scene S, const vx_camera* __restrict__ cam, FrameBuffers fb, uint32_t tiles_x,
                                                uint32_t ntiles, unsigned long long* __restrict__ counters) {
  const uint32_t tile = xcd_tile(blockIdx.x, ntiles);
  const uint32_t gx = (tile % tiles_x) * 8u + (threadIdx.x & 7u), gy = (tile / tiles_x) * 8u + (threadIdx.x >> 3);
  if constexpr (COUNT) {
    counted::Ctx C;
    C.S = S;
    C.cam = cam;
    C.grass_taper = nullptr;  // the godray pass has no grass geometry
    Counts cnt = {};
    C.cnt = &cnt;
    if (gx < fb.rw && gy < fb.rh) counted::godray_pixel(C, fb, gx, gy);
    flush_counts<true>(cnt, counters);
  } else {
    plain::Ctx C;
    C.S = S;
    C.cam = cam;
    C.grass_taper = nullptr;  // the godray pass has no grass geometry
    if (gx < fb.rw && gy < fb.rh) plain::godray_pixel(C, fb, gx, gy);
  }
}

__global__ __launch_bounds__(256) void k_composite(Scene S, const vx_camera* __restrict__ cam, FrameBuffers fb) {
  const uint32_t x = blockIdx.x * 16u + (threadIdx.x & 15u), y = blockIdx.y * 16u + (threadIdx.x >> 4);
  if (x >= fb.pw || y >= fb.ph) return;
  plain::Ctx C;
  C.S = S;
  C.cam = cam;
  C.grass_taper = nullptr;
  fb.pre_taa[(size_t)y * fb.pw + x] = plain::composite_pixel(C, fb, x, y);
}

__global__ __launch_bounds__(256) void k_composite_taa(Scene S, const vx_camera* __restrict__ cam, FrameBuffers fb) {
  const uint32_t x = blockIdx.x * 16u + (threadIdx.x & 15u), y = blockIdx.y * 16u + (threadIdx.x >> 4);
  if (x >= fb.pw || y >= fb.ph) return;
  plain::Ctx C;
  C.S = S;
  C.cam = cam;
  C.grass_taper = nullptr;
  plain::composite_taa_pixel(C, fb, x, y);
}

}  // namespace

// ---------------------------------------------------------------------------
// Launchers (host)
// ---------------------------------------------------------------------------

hipError_t vx_launch_primary(const vxd::Scene& S, const vx_camera* cam_dev, const vxd::FrameBuffers& fb, bool count,
                             unsigned long long* counters, hipStream_t st) {
  const uint32_t tx = (fb.rw + 7u) / 8u, ty = (fb.rh + 7u) / 8u, nt = tx * ty;
  if (count) hipLaunchKernelGGL(k_primary<true>, dim3(nt), dim3(64), 0, st, S, cam_dev, fb, tx, nt, counters);
  else hipLaunchKernelGGL(k_primary<false>, dim3(nt), dim3(64), 0, st, S, cam_dev, fb, tx, nt, counters);
  return hipGetLastError();
}
hipError_t vx_launch_local_taa(const vxd::FrameBuffers& fb, hipStream_t st) {
  hipLaunchKernelGGL(k_local_taa, dim3((fb.rw + 15u) / 16u, (fb.rh + 15u) / 16u), dim3(256), 0, st, fb);
  return hipGetLastError();
}
hipError_t vx_launch_godray(const vxd::Scene& S, const vx_camera* cam_dev, const vxd::FrameBuffers& fb, bool count,
                            unsigned long long* counters, hipStream_t st) {
  const uint32_t tx = (fb.rw + 7u) / 8u, ty = (fb.rh + 7u) / 8u, nt = tx * ty;
  if (count) hipLaunchKernelGGL(k_godray<true>, dim3(nt), dim3(64), 0, st, S, cam_dev, fb, tx, nt, counters);
  else hipLaunchKernelGGL(k_godray<false>, dim3(nt), dim3(64), 0, st, S, cam_dev, fb, tx, nt, counters);
  return hipGetLastError();
}
hipError_t vx_launch_composite(const vxd::Scene& S, const vx_camera* cam_dev, const vxd::FrameBuffers& fb, hipStream_t st) {
  hipLaunchKernelGGL(k_composite, dim3((fb.pw + 15u) / 16u, (fb.ph + 15u) / 16u), dim3(256), 0, st, S, cam_dev, fb);
  return hipGetLastError();
}
hipError_t vx_launch_composite_taa(const vxd::Scene& S, const vx_camera* cam_dev, const vxd::FrameBuffers& fb, hipStream_t st) {
  hipLaunchKernelGGL(k_composite_taa, dim3((fb.pw + 15u) / 16u, (fb.ph + 15u) / 16u), dim3(256), 0, st, S, cam_dev, fb);
  return hipGetLastError();
}
