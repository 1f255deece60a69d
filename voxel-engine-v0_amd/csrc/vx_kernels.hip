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

#include <algorithm>

#include "vx_launch.h"
#include "vx_terrain.h"
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
namespace executed {
#define VXD_COUNT 2
#include "vxd_passes.inc"
#undef VXD_COUNT
}  // namespace executed
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

// Bytes of a lane's counted loads and stores: element counts times the element
// sizes (grid word, ChunkMetaGpu, macro word, NodeGpu, NodeRopesGpu, colinfo
// word, clip texel; texture traffic is counted in bytes).
__device__ __forceinline__ uint32_t count_bytes(const Counts& c) {
  return 4u * c.grid + 48u * c.meta + 4u * c.macro + 16u * c.node + 32u * c.rope + 4u * c.colinfo + 4u * c.clip +
         c.tex_rd + c.tex_wr;
}

// Adds the wavefront's counts into out: the kNumCounts array counts, then
// the kNumParts part byte totals.
template <bool COUNT>
__device__ __forceinline__ void flush_counts(const Counts& c, unsigned long long* out) {
  if (!COUNT) return;
  const uint32_t v[kNumCounts] = {c.grid, c.meta, c.macro, c.node, c.rope, c.colinfo, c.clip, c.tex_rd, c.tex_wr};
#pragma unroll
  for (int i = 0; i < kNumCounts + (int)kNumParts; ++i) {
    unsigned long long s = i < kNumCounts ? v[i] : c.part_bytes[i - kNumCounts];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63u) == 0u) atomicAdd(out + i, s);
  }
}

// Occupancy of the primary-pass kernels: 4 waves per SIMD gives the compiler
// 128 VGPRs.  The pass needs more than that at its peak (the grass sphere
// trace inside the rope traversal), so it still spills; timed at 3840x2160
// the persistent kernel runs 15.8 ms at 4 waves and 16.6 at 3 (the LDS of a
// tile allows no more than 4): fewer spills win until latency hiding runs out.
#define VXD_PRIMARY_OCCUPANCY __attribute__((amdgpu_waves_per_eu(4, 8)))
// The godray kernels: 6 waves per SIMD (80 VGPRs, a few spills) beat 4
// (120 VGPRs, none) and 8 at 3840x2160, 3.57 against 3.90 and 3.62 ms: their
// sun-shadow traversals are latency-bound.
#define VXD_GODRAY_OCCUPANCY __attribute__((amdgpu_waves_per_eu(6, 8)))

// ---------------------------------------------------------------------------
// Primary pass, counting forms: one 8x8 tile per wavefront runs the whole
// pass per pixel with per-lane load counters.  k_primary_count_reference
// counts the reference algorithm's loads (the product's exact work skips
// off); k_primary_count_executed counts the loads the product kernel makes.
// ray_bytes receives each ray's bytes, the tile's gather shared over its 64 rays.
// ---------------------------------------------------------------------------
#define VXD_COUNTING_PRIMARY(NS, NAME)                                                                                  \
  __global__ __launch_bounds__(64) VXD_PRIMARY_OCCUPANCY void NAME(                                                    \
      Scene S, const vx_camera* __restrict__ cam, FrameBuffers fb, uint32_t tiles_x, uint32_t ntiles,                 \
      unsigned long long* __restrict__ counters, float* __restrict__ ray_bytes) {                                      \
    __shared__ TileLDS L;                                                                                              \
    const uint32_t tile = xcd_tile(blockIdx.x, ntiles);                                                                \
    const uint32_t wgx = tile % tiles_x, wgy = tile / tiles_x;                                                         \
    if (threadIdx.x < 12u) L.grass_taper[threadIdx.x] = plain::grass_taper_entry(threadIdx.x);                        \
    const uint32_t gx = wgx * 8u + (threadIdx.x & 7u), gy = wgy * 8u + (threadIdx.x >> 3);                             \
    NS::Ctx C;                                                                                                         \
    C.S = S;                                                                                                           \
    C.cam = cam;                                                                                                       \
    C.grass_taper = L.grass_taper;                                                                                     \
    C.grass_blades = L.grass_blades + threadIdx.x;                                                                     \
    Counts cnt = {};                                                                                                   \
    C.cnt = &cnt;                                                                                                      \
    if (threadIdx.x == 0) NS::tile_build(C, L, wgx, wgy);                                                              \
    const uint32_t tile_b = threadIdx.x == 0 ? count_bytes(cnt) : 0u; /* lane 0: the tile's gather so far */          \
    __syncthreads();                                                                                                   \
    if (gx < fb.rw && gy < fb.rh) NS::primary_pixel(C, L, fb, gx, gy, true, NS::sky_bg(F3(0.0f, 1.0f, 0.0f)));         \
    const float share = (float)__shfl(tile_b, 0, 64) * (1.0f / 64.0f);                                                 \
    if (ray_bytes && gx < fb.rw && gy < fb.rh)                                                                         \
      ray_bytes[(size_t)gy * fb.rw + gx] = (float)(count_bytes(cnt) - tile_b) + share;                                 \
    flush_counts<true>(cnt, counters);                                                                                 \
  }
VXD_COUNTING_PRIMARY(counted, k_primary_count_reference)
VXD_COUNTING_PRIMARY(executed, k_primary_count_executed)
#undef VXD_COUNTING_PRIMARY

__global__ __launch_bounds__(256) void k_local_taa(FrameBuffers fb) {
  const uint32_t x = blockIdx.x * 16u + (threadIdx.x & 15u), y = blockIdx.y * 16u + (threadIdx.x >> 4);
  if (x < fb.rw && y < fb.rh) plain::local_taa_pixel(fb, x, y);
}

template <bool COUNT>
__global__ __launch_bounds__(64) VXD_GODRAY_OCCUPANCY void k_godray(Scene S, const vx_camera* __restrict__ cam, FrameBuffers fb, uint32_t tiles_x,
                                                uint32_t ntiles, unsigned long long* __restrict__ counters) {
  const uint32_t tile = xcd_tile(blockIdx.x, ntiles);
  const uint32_t gx = (tile % tiles_x) * 8u + (threadIdx.x & 7u), gy = (tile / tiles_x) * 8u + (threadIdx.x >> 3);
  if constexpr (COUNT) {
    counted::Ctx C;
    C.S = S;
    C.cam = cam;
    C.grass_taper = nullptr;  // the godray pass has no grass geometry
    Counts cnt = {};
    cnt.part = kPartShadow;  // the godray pass's scene loads are its sun-shadow traces
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

// ---------------------------------------------------------------------------
// Ray queries: one persistent work-stealing launch over all rays.  Each lane
// owns one ray and advances it one chunk per loop trip; lanes whose ray has
// finished are refilled with new rays (one wavefront-aggregated atomic on the
// ray cursor) whenever fewer than `refill_below` lanes are still busy, so a
// wavefront does not idle behind its longest ray.  Workgroups exit when the
// ray cursor has passed n and none of their lanes holds a ray.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }

__global__ __launch_bounds__(64) void k_query_begin(uint32_t* __restrict__ cursor) {
  if (threadIdx.x == 0) cursor[0] = 0u;
}

// The persistent lane-refill loop over rays [0, n) from `cursor` (shared by
// the ray-query kernel and the fused batch launch), taken in the order `order`
// gives (null: input order).  Returns when the cursor has passed n and no lane
// of the wavefront holds a ray.
__device__ __forceinline__ void query_persistent(const plain::Ctx& C, const vx_ray* __restrict__ rays,
                                                 vx_hit* __restrict__ hits, uint32_t n, uint32_t* __restrict__ cursor,
                                                 uint32_t refill_below, const uint32_t* __restrict__ order = nullptr) {
  plain::QueryRay q;
  bool active = false, drained = false;
  uint32_t id = 0u;
  const unsigned long long below = (1ull << lane_id()) - 1ull;
  for (;;) {
    const uint32_t busy = (uint32_t)__popcll(__ballot(active));
    if (busy < refill_below) {
      const bool need = !active && !drained;
      const unsigned long long m = __ballot(need);
      if (m != 0ull) {
        uint32_t base = 0u;
        if (lane_id() == 0u) base = atomicAdd(cursor, (uint32_t)__popcll(m));
        base = __shfl(base, 0, 64);
        if (need) {
          id = base + (uint32_t)__popcll(m & below);
          if (id < n) {
            if (order) id = order[id];
            const vx_ray r = rays[id];
            if (plain::query_begin(C, F3(r.ox, r.oy, r.oz), F3(r.dx, r.dy, r.dz), r.tmax, q)) active = true;
            else hits[id] = plain::query_result(plain::miss_hitgeom());
          } else {
            drained = true;
          }
        }
      }
    }
    if (__ballot(active) == 0ull) {
      if (__ballot(!drained) == 0ull) break;
      continue;
    }
    if (active) {
      plain::HitGeom h;
      if (plain::query_step(C, q, h)) {
        hits[id] = plain::query_result(h);
        active = false;
      }
    }
  }
}

__global__ __launch_bounds__(64) VXD_PRIMARY_OCCUPANCY void k_query(
    Scene S, const vx_camera* __restrict__ cam, const vx_ray* __restrict__ rays, vx_hit* __restrict__ hits, uint32_t n,
    uint32_t* __restrict__ cursor, uint32_t refill_below, const uint32_t* __restrict__ order) {
  plain::Ctx C;
  C.S = S;
  C.cam = cam;
  C.grass_taper = nullptr;
  C.geom_only = true;
  query_persistent(C, rays, hits, n, cursor, refill_below, order);
}

// Sort keys of ray queries.  Incoherent rays traced in input order put rays
// from all over the world, in all directions, into one wavefront; sorted by
// this key, a wavefront's rays start in nearby space and point the same way,
// so they walk the same chunks and macro cells and load the same nodes.  The
// key only orders the work: each ray's result does not depend on it.
__device__ __forceinline__ uint32_t spread3(uint32_t v) {  // bits b of v to bit 3b
  uint32_t r = 0u;
#pragma unroll
  for (uint32_t b = 0; b < 5u; ++b) r |= ((v >> b) & 1u) << (3u * b);
  return r;
}
__global__ __launch_bounds__(256) void k_query_keys(const vx_camera* __restrict__ cam, const vx_ray* __restrict__ rays,
                                                    uint32_t n, uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= n) return;
  const vx_ray r = rays[i];
  const float csm = (float)cam->chunk_size * cam->voxel_params[0];
  const float o[3] = {r.ox, r.oy, r.oz}, d[3] = {r.dx, r.dy, r.dz};
  uint32_t qo[3], qd[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float lo = (float)cam->grid_origin_chunk[a] * csm, ext = (float)cam->grid_dims[a] * csm;
    // saturating conversions: NaN and out-of-box values land in the end bins
    qo[a] = (uint32_t)min(max(f2i((o[a] - lo) / ext * 32.0f), 0), 31);
    qd[a] = (uint32_t)min(max(f2i((d[a] + 1.0f) * 4.0f), 0), 7);
  }
  keys[i] = ((spread3(qd[0]) | (spread3(qd[1]) << 1u) | (spread3(qd[2]) << 2u)) << 15u) |
            spread3(qo[0]) | (spread3(qo[1]) << 1u) | (spread3(qo[2]) << 2u);
  vals[i] = i;
}

// Ray queries, counting forms: one ray per lane, traced alone (no refill) with
// per-lane load counters, k_query_count_reference counting the reference
// algorithm's loads and k_query_count_executed the loads k_query makes.  The
// traversal's loads are charged to the trace part; the ray record read and the
// hit record written count as texture bytes.
#define VXD_COUNTING_QUERY(NS, NAME)                                                                                    \
  __global__ __launch_bounds__(64) VXD_PRIMARY_OCCUPANCY void NAME(                                                    \
      Scene S, const vx_camera* __restrict__ cam, const vx_ray* __restrict__ rays, vx_hit* __restrict__ hits,          \
      uint32_t n, unsigned long long* __restrict__ counters) {                                                         \
    NS::Ctx C;                                                                                                         \
    C.S = S;                                                                                                           \
    C.cam = cam;                                                                                                       \
    C.grass_taper = nullptr;                                                                                           \
    C.geom_only = true;                                                                                                \
    Counts cnt = {};                                                                                                   \
    cnt.part = kPartTrace;                                                                                             \
    C.cnt = &cnt;                                                                                                      \
    const uint32_t id = blockIdx.x * 64u + threadIdx.x;                                                                \
    if (id < n) {                                                                                                      \
      cnt.tex_rd += (uint32_t)sizeof(vx_ray);                                                                          \
      cnt.part_bytes[kPartTextures] += (uint32_t)(sizeof(vx_ray) + sizeof(vx_hit));                                    \
      const vx_ray r = rays[id];                                                                                       \
      NS::QueryRay q;                                                                                                  \
      NS::HitGeom h = NS::miss_hitgeom();                                                                              \
      if (NS::query_begin(C, F3(r.ox, r.oy, r.oz), F3(r.dx, r.dy, r.dz), r.tmax, q))                                   \
        while (!NS::query_step(C, q, h)) {                                                                             \
        }                                                                                                              \
      hits[id] = NS::query_result(h);                                                                                  \
      cnt.tex_wr += (uint32_t)sizeof(vx_hit);                                                                          \
    }                                                                                                                  \
    flush_counts<true>(cnt, counters);                                                                                 \
  }
VXD_COUNTING_QUERY(counted, k_query_count_reference)
VXD_COUNTING_QUERY(executed, k_query_count_executed)
#undef VXD_COUNTING_QUERY

// ---------------------------------------------------------------------------
// Fused multi-view launches (a batch of views rendered together).
//
// Primary and godray run as persistent work-stealing kernels over the union
// of the views' 8x8 tiles (views in table order, largest first).  The tile
// space is cut into 8 contiguous bands, one per XCD; a workgroup takes tiles
// from its own XCD's band first (L2 locality), then steals from the other
// bands in turn.  Every workgroup exits once all 8 bands are exhausted.  The
// streaming passes run one flattened grid over all views' pixel blocks.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t batch_view_of(const Batch& B, uint32_t item, int which) {
  uint32_t lo = 0u, hi = B.n - 1u;  // last view whose base <= item
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1u) >> 1;
    const BatchView& V = B.views[mid];
    const uint32_t base = which == 0 ? V.tile_base
                        : which == 1 ? V.rblk_base
                        : which == 2 ? V.pblk_base
                                     : V.xt_base[which - 3];  // 3, 4: execution tiles of shapes 1, 2
    if (base <= item) lo = mid;
    else hi = mid - 1u;
  }
  return lo;
}

// Next run of up to B.grab of the `total` tiles for this workgroup, or false
// when every band is exhausted.  `band` is the band this workgroup is currently draining; it
// starts at the workgroup's XCD and advances when the band runs dry.
__device__ __forceinline__ bool batch_take(const Batch& B, uint32_t total, uint32_t pass, uint32_t& band, uint32_t& steps,
                                           uint32_t& t0, uint32_t& t1) {
  const uint32_t per = (total + kBatchBands - 1u) / kBatchBands;
  while (steps < kBatchBands) {
    const uint32_t lo = band * per, hi = min(total, lo + per);
    uint32_t got = 0u;
    if (threadIdx.x == 0) got = atomicAdd(B.cursor + pass * kBatchBands + band, B.grab);
    got = __builtin_amdgcn_readfirstlane(got);
    if (lo + got < hi) {
      t0 = lo + got;
      t1 = min(hi, t0 + B.grab);
      return true;
    }
    band = (band + 1u) % kBatchBands;
    ++steps;
  }
  return false;
}

__global__ __launch_bounds__(64) void k_batch_begin(Batch B) {
  if (threadIdx.x < 2u * kBatchBands) B.cursor[threadIdx.x] = 0u;
  if (B.rays && threadIdx.x == 0u) B.rays->cursor = 0u;
}

__global__ __launch_bounds__(64) void k_batch_end(Batch B) {
  for (uint32_t v = threadIdx.x; v < B.n; v += 64u) B.ping[v] ^= 1u;
}

// The 8x8 tile's candidate chunk list into L over the whole wavefront (the
// phases of plain::gather_*), or on one lane when a probe ray visits more
// cells than the phases hold.  The workgroup is one wavefront, and every
// branch below is uniform over it.
__device__ __forceinline__ void tile_gather(const plain::Ctx& C, TileLDS& L, uint32_t wgx, uint32_t wgy) {
  plain::gather_walk(C, L, wgx, wgy, threadIdx.x);
  __syncthreads();
  if (plain::gather_overflow(L)) {
    if (threadIdx.x == 0) plain::tile_build(C, L, wgx, wgy);
  } else {
    plain::gather_load(C, L, threadIdx.x);
    __syncthreads();
    plain::gather_first(L, threadIdx.x);
    __syncthreads();
    plain::gather_place(L, threadIdx.x);
    __syncthreads();
    if (threadIdx.x == 0) plain::gather_finish(L);
  }
  __syncthreads();
}

static_assert(PRIMARY_MAX_TILE_CHUNKS + 1u <= kListWords, "a list holds its count and the cached slots");

// The candidate lists of every 8x8 tile of the batch's views, for the
// execution tile shapes other than 8x8: one workgroup per tile (XCD-banded),
// the same gather as the primary kernel's, stored as [count, slots...].
__global__ __launch_bounds__(64) void k_tile_lists(Scene S, Batch B) {
  __shared__ TileLDS L;
  const uint32_t t = xcd_tile(blockIdx.x, B.tiles);
  const uint32_t v = batch_view_of(B, t, 0);
  const BatchView& V = B.views[v];
  const uint32_t tile = t - V.tile_base;
  plain::Ctx C;
  C.S = S;
  C.cam = V.cam;
  C.grass_taper = nullptr;
  tile_gather(C, L, tile % V.tiles_x, tile / V.tiles_x);
  uint32_t* out = B.lists + (size_t)t * kListWords;
  const uint32_t n = L.cached;
  if (threadIdx.x == 0) out[0] = n;
  if (threadIdx.x < n) out[1u + threadIdx.x] = L.slots[threadIdx.x];
}

// PROF: the launch carries a view whose camera has profile_flags & 1; the
// product launch (PROF = false) has no profiling code at all.
//
// SHAPE: the execution tile (0: 8x8, 1: 16x4, 2: 32x2).  With 8x8 the
// wavefront builds its tile's candidate list in LDS; with the others each lane
// traces with its own 8x8 tile's list from the pre-pass (k_tile_lists).
template <bool PROF, uint32_t SHAPE>
__global__ __launch_bounds__(64) VXD_PRIMARY_OCCUPANCY void k_primary_batch(Scene S, Batch B) {
  constexpr uint32_t TW = 8u << SHAPE, TH = 64u / TW;
  __shared__ TileLDS L;
  // One deferral region per workgroup (the launchers never start more
  // workgroups than there are regions).
  if (blockIdx.x >= S.n_defer_regions) return;
  uint4* const defer = S.defer + (size_t)blockIdx.x * (kDeferEntries * 64u) + threadIdx.x;
  if (threadIdx.x < 12u) L.grass_taper[threadIdx.x] = plain::grass_taper_entry(threadIdx.x);
  if (threadIdx.x < 5u) L.prof[threadIdx.x] = 0u;
  const uint32_t total = SHAPE == 0u ? B.tiles : B.xtiles[SHAPE - 1u];
  uint32_t band = blockIdx.x % kBatchBands, steps = 0u, t0 = 0u, t1 = 0u;
  while (batch_take(B, total, 0u, band, steps, t0, t1)) {
    for (uint32_t t = t0; t < t1; ++t) {
      const uint32_t v = batch_view_of(B, t, SHAPE == 0u ? 0 : (int)(2u + SHAPE));
      const BatchView& V = B.views[v];
      const FrameBuffers& fb = V.fb[B.ping[v] & 1u];
      plain::Ctx C;
      C.S = S;
      C.cam = V.cam;
      C.grass_taper = L.grass_taper;
      C.grass_blades = L.grass_blades + threadIdx.x;
      C.defer = defer;
      // Profiling (ray_main.wgsl:64-74) is a property of the view's camera, so
      // uniform over the tile: the counters go to LDS and are flushed per tile.
      const bool prof = PROF && (V.cam->profile_flags & 1u) != 0u;
      C.prof = prof ? L.prof : nullptr;
      __syncthreads();  // the previous tile's lanes are done with L
      // A view's tiles are taken bottom row first: the rows near the camera's
      // feet (grass close up) cost the most and the sky rows the least, so a
      // band's last tiles, and with them the launch's tail, are cheap ones.
      // The order changes no result (each tile is traced on its own).
      if constexpr (SHAPE == 0u) {
        const uint32_t tile = V.tiles - 1u - (t - V.tile_base);
        const uint32_t wgx = tile % V.tiles_x, wgy = tile / V.tiles_x;
        tile_gather(C, L, wgx, wgy);
        // Every lane runs the pixel pass (those outside the image only help
        // with the deferred grass tests, which need the whole wavefront).
        const uint32_t gx = wgx * 8u + (threadIdx.x & 7u), gy = wgy * 8u + (threadIdx.x >> 3);
        plain::primary_pixel(C, plain::LdsList{L}, fb, gx, gy, gx < fb.rw && gy < fb.rh, plain::sky_bg(F3(0.0f, 1.0f, 0.0f)));
      } else {
        const uint32_t xtx = V.xt_x[SHAPE - 1u];
        const uint32_t xt = xtx * ((fb.rh + TH - 1u) / TH) - 1u - (t - V.xt_base[SHAPE - 1u]);
        const uint32_t gx = (xt % xtx) * TW + (threadIdx.x % TW), gy = (xt / xtx) * TH + (threadIdx.x / TW);
        // this pixel's 8x8 tile (clamped into the view for lanes past its edge)
        const uint32_t ty = V.tiles / V.tiles_x;
        const uint32_t lt = min(gy >> 3, ty - 1u) * V.tiles_x + min(gx >> 3, V.tiles_x - 1u);
        const plain::LaneList lst{B.lists + (size_t)(V.tile_base + lt) * kListWords};
        plain::primary_pixel(C, lst, fb, gx, gy, gx < fb.rw && gy < fb.rh, plain::sky_bg(F3(0.0f, 1.0f, 0.0f)));
      }
      if (prof) {
        __syncthreads();  // every lane's LDS increments are in
        if (threadIdx.x < 5u) {
          const uint32_t v = L.prof[threadIdx.x];
          if (v) atomicAdd(fb.prof + threadIdx.x, v);
          L.prof[threadIdx.x] = 0u;
        }
      }
    }
  }
  // Every band is exhausted: the launch's ray queries fill the rest of the
  // workgroup's life (the same lane-refill loop as k_query), so workgroups
  // that finish their tiles early trace rays while the last tiles finish.
  if (B.rays) {
    BatchRays* const R = B.rays;
    const uint32_t n = __builtin_amdgcn_readfirstlane(R->n);
    if (n) {
      plain::Ctx C;
      C.S = S;
      C.cam = &R->cam;
      C.grass_taper = nullptr;
      C.geom_only = true;
      query_persistent(C, R->rays, R->hits, n, &R->cursor, __builtin_amdgcn_readfirstlane(R->refill_below), R->order);
    }
  }
}

// Godray: one workgroup per tile of every view (a flattened grid; the pass is
// short and even, so plain dispatch beats work stealing), XCD bands as in k_godray.
__global__ __launch_bounds__(64) VXD_GODRAY_OCCUPANCY void k_godray_batch(Scene S, Batch B) {
  const uint32_t t = xcd_tile(blockIdx.x, B.tiles);
  const uint32_t v = batch_view_of(B, t, 0);
  const BatchView& V = B.views[v];
  const FrameBuffers& fb = V.fb[B.ping[v] & 1u];
  const uint32_t tile = t - V.tile_base;
  const uint32_t gx = (tile % V.tiles_x) * 8u + (threadIdx.x & 7u), gy = (tile / V.tiles_x) * 8u + (threadIdx.x >> 3);
  plain::Ctx C;
  C.S = S;
  C.cam = V.cam;
  C.grass_taper = nullptr;
  if (gx < fb.rw && gy < fb.rh) plain::godray_pixel(C, fb, gx, gy);
}

__global__ __launch_bounds__(256) void k_local_taa_batch(Batch B) {
  const uint32_t v = batch_view_of(B, blockIdx.x, 1);
  const BatchView& V = B.views[v];
  const FrameBuffers& fb = V.fb[B.ping[v] & 1u];
  const uint32_t blk = blockIdx.x - V.rblk_base;
  const uint32_t x = (blk % V.rblk_x) * 16u + (threadIdx.x & 15u), y = (blk / V.rblk_x) * 16u + (threadIdx.x >> 4);
  if (x < fb.rw && y < fb.rh) plain::local_taa_pixel(fb, x, y);
}

__global__ __launch_bounds__(256) void k_composite_batch(Scene S, Batch B) {
  const uint32_t v = batch_view_of(B, blockIdx.x, 2);
  const BatchView& V = B.views[v];
  const FrameBuffers& fb = V.fb[B.ping[v] & 1u];
  const uint32_t blk = blockIdx.x - V.pblk_base;
  const uint32_t x = (blk % V.pblk_x) * 16u + (threadIdx.x & 15u), y = (blk / V.pblk_x) * 16u + (threadIdx.x >> 4);
  if (x >= fb.pw || y >= fb.ph) return;
  plain::Ctx C;
  C.S = S;
  C.cam = V.cam;
  C.grass_taper = nullptr;
  fb.pre_taa[(size_t)y * fb.pw + x] = plain::composite_pixel(C, fb, x, y);
}

__global__ __launch_bounds__(256) void k_composite_taa_batch(Scene S, Batch B) {
  const uint32_t v = batch_view_of(B, blockIdx.x, 2);
  const BatchView& V = B.views[v];
  const FrameBuffers& fb = V.fb[B.ping[v] & 1u];
  const uint32_t blk = blockIdx.x - V.pblk_base;
  const uint32_t x = (blk % V.pblk_x) * 16u + (threadIdx.x & 15u), y = (blk / V.pblk_x) * 16u + (threadIdx.x >> 4);
  if (x >= fb.pw || y >= fb.ph) return;
  plain::Ctx C;
  C.S = S;
  C.cam = V.cam;
  C.grass_taper = nullptr;
  plain::composite_taa_pixel(C, fb, x, y);
}

}  // namespace

// ---------------------------------------------------------------------------
// Launchers (host)
// ---------------------------------------------------------------------------

hipError_t vx_launch_primary_count(const vxd::Scene& S, const vx_camera* cam_dev, const vxd::FrameBuffers& fb,
                                   bool executed, unsigned long long* counters, float* ray_bytes, hipStream_t st) {
  const uint32_t tx = (fb.rw + 7u) / 8u, ty = (fb.rh + 7u) / 8u, nt = tx * ty;
  if (executed)
    hipLaunchKernelGGL(k_primary_count_executed, dim3(nt), dim3(64), 0, st, S, cam_dev, fb, tx, nt, counters, ray_bytes);
  else
    hipLaunchKernelGGL(k_primary_count_reference, dim3(nt), dim3(64), 0, st, S, cam_dev, fb, tx, nt, counters, ray_bytes);
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

// Fused passes of a batch in reference order; `persistent` is the number of
// workgroups of the two work-stealing kernels.
hipError_t vx_launch_batch(const vxd::Scene& S, const vxd::Scene& S_godray, const vxd::Batch& B, uint32_t mask,
                           uint32_t persistent, bool profile, hipStream_t st) {
  hipLaunchKernelGGL(k_batch_begin, dim3(1), dim3(64), 0, st, B);
  if (mask & VX_PASS_PRIMARY) {
    const uint32_t groups = std::min(persistent, S.n_defer_regions);
    if (B.shape != 0u && B.tiles) hipLaunchKernelGGL(k_tile_lists, dim3(B.tiles), dim3(64), 0, st, S, B);
    if (B.shape == 1u) {
      if (profile) hipLaunchKernelGGL((k_primary_batch<true, 1u>), dim3(groups), dim3(64), 0, st, S, B);
      else hipLaunchKernelGGL((k_primary_batch<false, 1u>), dim3(groups), dim3(64), 0, st, S, B);
    } else if (B.shape == 2u) {
      if (profile) hipLaunchKernelGGL((k_primary_batch<true, 2u>), dim3(groups), dim3(64), 0, st, S, B);
      else hipLaunchKernelGGL((k_primary_batch<false, 2u>), dim3(groups), dim3(64), 0, st, S, B);
    } else {
      if (profile) hipLaunchKernelGGL((k_primary_batch<true, 0u>), dim3(groups), dim3(64), 0, st, S, B);
      else hipLaunchKernelGGL((k_primary_batch<false, 0u>), dim3(groups), dim3(64), 0, st, S, B);
    }
  }
  if (mask & VX_PASS_LOCAL_TAA) hipLaunchKernelGGL(k_local_taa_batch, dim3(B.rblks), dim3(256), 0, st, B);
  if (mask & VX_PASS_GODRAY) hipLaunchKernelGGL(k_godray_batch, dim3(B.tiles), dim3(64), 0, st, S_godray, B);
  if (mask & VX_PASS_COMPOSITE) hipLaunchKernelGGL(k_composite_batch, dim3(B.pblks), dim3(256), 0, st, S, B);
  if (mask & VX_PASS_COMPOSITE_TAA) hipLaunchKernelGGL(k_composite_taa_batch, dim3(B.pblks), dim3(256), 0, st, S, B);
  if (mask & VX_PASS_END_FRAME) hipLaunchKernelGGL(k_batch_end, dim3(1), dim3(64), 0, st, B);
  return hipGetLastError();
}

// One view through the work-stealing kernels (vx_render's default form).
hipError_t vx_launch_ws_begin(const vxd::Batch& B, hipStream_t st) {
  hipLaunchKernelGGL(k_batch_begin, dim3(1), dim3(64), 0, st, B);
  return hipGetLastError();
}
hipError_t vx_launch_primary_ws(const vxd::Scene& S, const vxd::Batch& B, uint32_t persistent, bool profile,
                                hipStream_t st) {
  const uint32_t groups = std::min(persistent, S.n_defer_regions);
  if (profile) hipLaunchKernelGGL((k_primary_batch<true, 0u>), dim3(groups), dim3(64), 0, st, S, B);
  else hipLaunchKernelGGL((k_primary_batch<false, 0u>), dim3(groups), dim3(64), 0, st, S, B);
  return hipGetLastError();
}

// Per chunk slot, whether its node range holds a LIGHT leaf (the flags the
// local-light gather tests before it marches).  One workgroup per slot; the
// range [node_base, node_base + node_count) is clipped to the arena.
__global__ __launch_bounds__(256) void k_chunk_lights(Scene S, uint32_t* __restrict__ out) {
  const uint32_t slot = blockIdx.x;
  const vx_chunk_meta& m = S.chunks[slot];
  const uint64_t lo = min((uint64_t)m.node_base, S.n_nodes), hi = min(lo + (uint64_t)m.node_count, S.n_nodes);
  int any = 0;
  for (uint64_t i = lo + threadIdx.x; i < hi; i += 256u) {
    const uint4 n = S.nodes[i];
    any |= (n.x == LEAF_U32 && n.z == MAT_LIGHT) ? 1 : 0;
  }
  any = __syncthreads_or(any);
  if (threadIdx.x == 0) out[slot] = any ? 1u : 0u;
}

hipError_t vx_launch_chunk_lights(const vxd::Scene& S, uint32_t* out, uint32_t n_slots, hipStream_t st) {
  if (n_slots) hipLaunchKernelGGL(k_chunk_lights, dim3(n_slots), dim3(256), 0, st, S, out);
  return hipGetLastError();
}

// Clipmap texels on the device: workgroup (bx, r) computes texels
// [256 bx, 256 bx + 256) of rectangle r in row-major order and stores each at
// its storage position; the value is the host producer's vxt::clip_texel_height.
// Rectangles of one update may overlap (a column strip and a row strip share a
// corner, and with the reference's strip placement they disagree there): as
// with uploads applied in order, a texel is written only by the last rectangle
// covering it.  Rectangles are checked against the array (level, extent, res)
// on the host.
__global__ __launch_bounds__(256) void k_clip_fill(const vx_clip_rect* __restrict__ rects, uint32_t n, vx_world_params wp,
                                                   float* __restrict__ clip_height) {
  const vx_clip_rect r = rects[blockIdx.y];
  const uint32_t t = blockIdx.x * 256u + threadIdx.x;
  if (t >= r.w * r.h) return;
  const uint32_t j = t / r.w, i = t - j * r.w;
  const uint32_t sx = r.x + i, sy = r.y + j;
  for (uint32_t q = blockIdx.y + 1; q < n; ++q) {
    const vx_clip_rect o = rects[q];
    if (o.level == r.level && sx - o.x < o.w && sy - o.y < o.h) return;  // unsigned: also rejects sx < o.x
  }
  const float h = vxt::clip_texel_height(wp, r.origin_x_m, r.origin_z_m, r.cell_m, (r.lx0 + i) % r.res, (r.lz0 + j) % r.res);
  clip_height[((size_t)r.level * r.res + sy) * r.res + sx] = h;
}

hipError_t vx_launch_clip_fill(const vx_clip_rect* rects, uint32_t n, uint32_t max_texels, const vx_world_params& world,
                               float* clip_height, hipStream_t st) {
  if (n && max_texels)
    hipLaunchKernelGGL(k_clip_fill, dim3((max_texels + 255u) / 256u, n), dim3(256), 0, st, rects, n, world, clip_height);
  return hipGetLastError();
}

// Ray queries: reset the cursor, then one persistent launch over all n rays.
hipError_t vx_launch_query(const vxd::Scene& S, const vx_camera* cam_dev, const vx_ray* rays, vx_hit* hits, uint32_t n,
                           uint32_t* cursor, uint32_t persistent, uint32_t refill_below, hipStream_t st) {
  hipLaunchKernelGGL(k_query_begin, dim3(1), dim3(64), 0, st, cursor);
  hipLaunchKernelGGL(k_query, dim3(persistent), dim3(64), 0, st, S, cam_dev, rays, hits, n, cursor, refill_below,
                     (const uint32_t*)nullptr);
  return hipGetLastError();
}
hipError_t vx_launch_query_ordered(const vxd::Scene& S, const vx_camera* cam_dev, const vx_ray* rays, vx_hit* hits,
                                   uint32_t n, const uint32_t* order, uint32_t* cursor, uint32_t persistent,
                                   uint32_t refill_below, hipStream_t st) {
  hipLaunchKernelGGL(k_query_begin, dim3(1), dim3(64), 0, st, cursor);
  hipLaunchKernelGGL(k_query, dim3(persistent), dim3(64), 0, st, S, cam_dev, rays, hits, n, cursor, refill_below, order);
  return hipGetLastError();
}
hipError_t vx_launch_query_keys(const vx_camera* cam_dev, const vx_ray* rays, uint32_t n, uint32_t* keys, uint32_t* vals,
                                hipStream_t st) {
  static_assert(kQuerySortBits == 24u, "9 direction bits above 15 origin bits");
  if (n) hipLaunchKernelGGL(k_query_keys, dim3((n + 255u) / 256u), dim3(256), 0, st, cam_dev, rays, n, keys, vals);
  return hipGetLastError();
}
hipError_t vx_launch_query_count(const vxd::Scene& S, const vx_camera* cam_dev, const vx_ray* rays, vx_hit* hits, uint32_t n,
                                 bool executed, unsigned long long* counters, hipStream_t st) {
  const uint32_t groups = (n + 63u) / 64u;
  if (!groups) return hipSuccess;
  if (executed)
    hipLaunchKernelGGL(k_query_count_executed, dim3(groups), dim3(64), 0, st, S, cam_dev, rays, hits, n, counters);
  else
    hipLaunchKernelGGL(k_query_count_reference, dim3(groups), dim3(64), 0, st, S, cam_dev, rays, hits, n, counters);
  return hipGetLastError();
}
