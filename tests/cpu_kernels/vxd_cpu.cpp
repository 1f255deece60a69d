// vxd_cpu.cpp — the product's device pass code (vxd_passes.inc) compiled for
// the host, with host loops standing in for the kernels' tile, lane and queue
// structure.  Test infrastructure only: it lets the CPU test suite check the
// exact code the GPU runs (results, load counts, loop termination) against
// the oracle without a GPU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

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
// The product form with 2 deferred grass tests per lane and attempt, so that
// chunks take repeated attempts (mode 3).
namespace plain_k2 {
#define VXD_COUNT 0
#define VXD_DEFER_K 2
#include "vxd_passes.inc"
#undef VXD_DEFER_K
#undef VXD_COUNT
}  // namespace plain_k2
}  // namespace vxd

// The lane the host loops run (each lane is a wavefront of one).
thread_local uint32_t vxd_cpu_lane = 0;

using namespace vxd;

extern "C" {

// Same layout as the oracle's scene description.
typedef struct cpu_scene {
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
} cpu_scene;

}  // extern "C"

namespace {

// The per-slot LIGHT flags the library derives on the device (k_chunk_lights):
// a slot is flagged when its node range holds a LIGHT leaf.
std::vector<uint32_t> chunk_lights(const cpu_scene* s) {
  std::vector<uint32_t> out(s->n_chunks, 0u);
  for (uint64_t slot = 0; slot < s->n_chunks; ++slot) {
    const vx_chunk_meta& m = s->chunks[slot];
    const uint64_t lo = std::min<uint64_t>(m.node_base, s->n_nodes), hi = std::min<uint64_t>(lo + m.node_count, s->n_nodes);
    for (uint64_t i = lo; i < hi && !out[slot]; ++i)
      if (s->nodes[i].child_base == LEAF_U32 && s->nodes[i].material == MAT_LIGHT) out[slot] = 1u;
  }
  return out;
}

Scene to_scene(const cpu_scene* s, const std::vector<uint32_t>* lights) {
  Scene S;
  S.chunks = s->chunks;
  S.nodes = reinterpret_cast<const uint4*>(s->nodes);
  S.ropes = reinterpret_cast<const uint4*>(s->ropes);
  S.grid = s->grid;
  S.macro = s->macro_words;
  S.colinfo = s->colinfo;
  S.clip = s->clip;
  S.clip_height = s->clip_height;
  S.n_chunks = s->n_chunks;
  S.n_nodes = s->n_nodes;
  S.n_grid = s->n_grid;
  S.n_macro = s->n_macro;
  S.n_colinfo = s->n_colinfo;
  S.n_clip = s->n_clip;
  S.chunk_light = lights ? lights->data() : nullptr;
  return S;
}

template <class F>
void parallel_for(uint32_t n, int threads, F f) {
  if (threads <= 1 || n < 2) {
    for (uint32_t i = 0; i < n; ++i) f(i, 0);
    return;
  }
  std::atomic<uint32_t> next{0};
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; ++t)
    pool.emplace_back([&, t] {
      for (uint32_t i; (i = next.fetch_add(1)) < n;) f(i, t);
    });
  for (auto& th : pool) th.join();
}

void fill_taper(float* taper) {
  for (uint32_t j = 0; j < 12u; ++j) taper[j] = plain::grass_taper_entry(j);
}

}  // namespace

extern "C" {

// mode 0: the product pass (k_primary_batch), which adds the profile counters
// into prof[5] when cam->profile_flags & 1 (as k_primary_batch does through
// each tile's LDS counters); 3: the same with 2 deferred grass tests per lane
// and attempt; 2: the load-counting kernel (k_primary_count_reference), which
// also adds its counts into counts[9] and its part byte totals into
// counts[9..17).
int vxd_cpu_primary(const cpu_scene* scene, const vx_camera* cam, uint32_t w, uint32_t h, const float* primary_hist_in,
                    const float* shadow_in, float* color, float* depth, float* local, float* primary_hist_out,
                    float* shadow_out, vx_hit* hits, int mode, int threads, uint64_t* counts, uint64_t* prof) {
  if (!scene || !cam || !w || !h) return -1;
  const std::vector<uint32_t> lights = chunk_lights(scene);
  const Scene S = to_scene(scene, &lights);
  FrameBuffers fb{};
  fb.rw = fb.pw = w;
  fb.rh = fb.ph = h;
  fb.primary_hist_in = reinterpret_cast<const float4*>(primary_hist_in);
  fb.primary_hist_out = reinterpret_cast<float4*>(primary_hist_out);
  fb.shadow_in = shadow_in;
  fb.shadow_out = shadow_out;
  fb.color = reinterpret_cast<float4*>(color);
  fb.depth = depth;
  fb.local = reinterpret_cast<float4*>(local);
  fb.hits = hits;
  const uint32_t tx = (w + 7u) / 8u, ty = (h + 7u) / 8u, nt = tx * ty;
  threads = std::max(1, threads);
  std::vector<Counts> cnt((size_t)threads, Counts{});
  std::vector<uint64_t> prof_t((size_t)threads * 5, 0u);

  // The product pass for one tile in namespace NS (its wavefront gather, then
  // each lane's pixel with the lane's own deferral column).
#define VXD_CPU_PRODUCT_TILE(NS)                                                                                     \
  do {                                                                                                               \
    NS::Ctx C;                                                                                                       \
    C.S = S;                                                                                                         \
    C.cam = cam;                                                                                                     \
    C.grass_taper = L.grass_taper;                                                                                   \
    for (uint32_t& v : L.prof) v = 0u;                                                                               \
    C.prof = (cam->profile_flags & 1u) ? L.prof : nullptr;                                                           \
    /* k_primary_batch's wavefront gather: each phase for every lane, the barrier between phases being the end of  \
       each lane loop */                                                                                             \
    for (uint32_t lane = 0; lane < 64u; ++lane) NS::gather_walk(C, L, wgx, wgy, lane);                               \
    if (NS::gather_overflow(L)) {                                                                                    \
      NS::tile_build(C, L, wgx, wgy);                                                                                \
    } else {                                                                                                         \
      for (uint32_t lane = 0; lane < 64u; ++lane) NS::gather_load(C, L, lane);                                       \
      for (uint32_t lane = 0; lane < 64u; ++lane) NS::gather_first(L, lane);                                         \
      for (uint32_t lane = 0; lane < 64u; ++lane) NS::gather_place(L, lane);                                         \
      NS::gather_finish(L);                                                                                          \
    }                                                                                                                \
    std::vector<uint4> defer((size_t)kDeferEntries * 64u);                                                           \
    for (uint32_t lane = 0; lane < 64u; ++lane) {                                                                    \
      const uint32_t gx = wgx * 8u + (lane & 7u), gy = wgy * 8u + (lane >> 3);                                       \
      vxd_cpu_lane = lane;                                                                                           \
      C.grass_blades = L.grass_blades + lane;                                                                        \
      C.defer = defer.data() + lane;                                                                                 \
      if (gx < w && gy < h) NS::primary_pixel(C, L, fb, gx, gy, true, NS::sky_bg(F3(0.0f, 1.0f, 0.0f)));             \
    }                                                                                                                \
    for (int i = 0; i < 5; ++i) prof_t[(size_t)t * 5 + i] += L.prof[i];                                              \
  } while (0)
  if (mode == 0 || mode == 2 || mode == 3) {
    parallel_for(nt, threads, [&](uint32_t tile, int t) {
      TileLDS L;
      fill_taper(L.grass_taper);
      const uint32_t wgx = tile % tx, wgy = tile / tx;
      if (mode == 0) {
        VXD_CPU_PRODUCT_TILE(plain);
      } else if (mode == 3) {
        VXD_CPU_PRODUCT_TILE(plain_k2);
      } else {
        counted::Ctx C;
        C.S = S;
        C.cam = cam;
        C.grass_taper = L.grass_taper;
        C.cnt = &cnt[t];
        counted::tile_build(C, L, wgx, wgy);
        for (uint32_t lane = 0; lane < 64u; ++lane) {
          const uint32_t gx = wgx * 8u + (lane & 7u), gy = wgy * 8u + (lane >> 3);
          C.grass_blades = L.grass_blades + lane;
          if (gx < w && gy < h) counted::primary_pixel(C, L, fb, gx, gy, true, counted::sky_bg(F3(0.0f, 1.0f, 0.0f)));
        }
      }
    });
  } else {
    return -1;
  }
#undef VXD_CPU_PRODUCT_TILE
  if (prof)
    for (int t = 0; t < threads; ++t)
      for (int i = 0; i < 5; ++i) prof[i] += prof_t[(size_t)t * 5 + i];
  if (counts) {
    for (const Counts& c : cnt) {
      const uint32_t v[kNumCounts] = {c.grid, c.meta, c.macro, c.node, c.rope, c.colinfo, c.clip, c.tex_rd, c.tex_wr};
      for (int i = 0; i < kNumCounts; ++i) counts[i] += v[i];
      for (int i = 0; i < (int)kNumParts; ++i) counts[kNumCounts + i] += c.part_bytes[i];
    }
  }
  return 0;
}

// The device's floor of a quotient (floor_quot with quot_rcp), elementwise.
int vxd_cpu_floor_quot(const float* x, const float* d, uint32_t n, float* out) {
  for (uint32_t k = 0; k < n; ++k) out[k] = plain::floor_quot(x[k], d[k], plain::quot_rcp(d[k]));
  return 0;
}

// The device float->int conversions, elementwise.
int vxd_cpu_convert(const float* x, uint32_t n, int32_t* out_i, uint32_t* out_u) {
  for (uint32_t k = 0; k < n; ++k) {
    out_i[k] = f2i(x[k]);
    out_u[k] = f2u(x[k]);
  }
  return 0;
}

// The device's constant-divisor quotient (divc<k>) against the correctly
// rounded division, over the non-negative binary32 bit patterns [lo, hi):
// returns the number of inputs where they differ (NaN results compare equal).
// Threads split the range.  Hardware fma for the two fmaf of divc.
extern "C++" {
template <uint32_t K>
__attribute__((target("fma"))) static uint64_t fastdiv_range(uint32_t lo, uint32_t hi) {
  constexpr float d = FAST_DIVISORS[K];
  uint64_t bad = 0;
  for (uint32_t b = lo; b < hi; ++b) {
    float x;
    std::memcpy(&x, &b, 4);
    const float f = divc<K>(x), r = x / d;
    uint32_t fb, rb;
    std::memcpy(&fb, &f, 4);
    std::memcpy(&rb, &r, 4);
    bad += (fb != rb) && !(f != f && r != r);
  }
  return bad;
}
}  // extern "C++"
uint64_t vxd_cpu_fastdiv_mismatches(uint32_t k, uint32_t lo, uint32_t hi, uint32_t threads) {
  static uint64_t (*const fns[])(uint32_t, uint32_t) = {fastdiv_range<0>, fastdiv_range<1>, fastdiv_range<2>,
                                                         fastdiv_range<3>, fastdiv_range<4>, fastdiv_range<5>};
  static_assert(sizeof(fns) / sizeof(fns[0]) == DIV_COUNT, "one checker per divisor");
  if (k >= DIV_COUNT || hi < lo) return ~0ull;
  threads = std::max(1u, threads);
  std::vector<uint64_t> bad(threads, 0);
  std::vector<std::thread> pool;
  const uint64_t span = (uint64_t)hi - lo, per = (span + threads - 1) / threads;
  for (uint32_t t = 0; t < threads; ++t) {
    const uint64_t a = lo + std::min<uint64_t>(span, t * per), b = lo + std::min<uint64_t>(span, (t + 1) * per);
    pool.emplace_back([&, t, a, b] { bad[t] = fns[k]((uint32_t)a, (uint32_t)b); });
  }
  for (auto& th : pool) th.join();
  uint64_t sum = 0;
  for (uint64_t v : bad) sum += v;
  return sum;
}
float vxd_cpu_fast_divisor(uint32_t k) { return k < DIV_COUNT ? FAST_DIVISORS[k] : 0.0f; }
uint32_t vxd_cpu_fast_divisor_count() { return DIV_COUNT; }

// The device value noise and fbm, elementwise: out[2n] = value_noise, fbm at (x, y).
int vxd_cpu_noise(const float* x, const float* y, uint32_t n, float* out) {
  for (uint32_t k = 0; k < n; ++k) {
    out[2 * k + 0] = plain::value_noise(F2(x[k], y[k]));
    out[2 * k + 1] = plain::fbm(F2(x[k], y[k]));
  }
  return 0;
}

// The device transcendentals, elementwise: out[5n] = sin, cos, exp, log, pow(x, y).
int vxd_cpu_math(const float* x, const float* y, uint32_t n, float* out) {
  for (uint32_t k = 0; k < n; ++k) {
    out[5 * k + 0] = vx_sin(x[k]);
    out[5 * k + 1] = vx_cos(x[k]);
    out[5 * k + 2] = vx_exp(x[k]);
    out[5 * k + 3] = vx_log(x[k]);
    out[5 * k + 4] = vx_pow(x[k], y[k]);
  }
  return 0;
}

// The tile candidate lists of a w x h view: mode 0 the one-lane gather
// (tile_build), mode 1 the wavefront phases of k_primary_batch.  Per tile:
// count (the raw counter), cached, and the first 24 slots and enter t.
int vxd_cpu_tile_lists(const cpu_scene* scene, const vx_camera* cam, uint32_t w, uint32_t h, int mode, uint32_t* counts,
                       uint32_t* cached, uint32_t* slots, float* enter, uint32_t* overflow) {
  if (!scene || !cam || !w || !h) return -1;
  const std::vector<uint32_t> lights = chunk_lights(scene);
  const Scene S = to_scene(scene, &lights);
  const uint32_t tx = (w + 7u) / 8u, ty = (h + 7u) / 8u;
  uint32_t over = 0;
  for (uint32_t tile = 0; tile < tx * ty; ++tile) {
    TileLDS L;
    std::memset(&L, 0xAB, sizeof(L));  // stale contents, as LDS holds between tiles
    plain::Ctx C;
    C.S = S;
    C.cam = cam;
    C.grass_taper = L.grass_taper;
    const uint32_t wgx = tile % tx, wgy = tile / tx;
    if (mode == 0) {
      plain::tile_build(C, L, wgx, wgy);
    } else {
      for (uint32_t lane = 0; lane < 64u; ++lane) plain::gather_walk(C, L, wgx, wgy, lane);
      if (plain::gather_overflow(L)) {
        ++over;
        plain::tile_build(C, L, wgx, wgy);
      } else {
        for (uint32_t lane = 0; lane < 64u; ++lane) plain::gather_load(C, L, lane);
        for (uint32_t lane = 0; lane < 64u; ++lane) plain::gather_first(L, lane);
        for (uint32_t lane = 0; lane < 64u; ++lane) plain::gather_place(L, lane);
        plain::gather_finish(L);
      }
    }
    counts[tile] = L.count;
    cached[tile] = L.cached;
    for (uint32_t i = 0; i < 24u; ++i) {
      slots[tile * 24u + i] = i < L.cached ? L.slots[i] : 0u;
      enter[tile * 24u + i] = i < L.cached ? L.enter[i] : 0.0f;
    }
  }
  if (overflow) *overflow = over;
  return 0;
}

// Ray queries: the device query code (begin + one-chunk steps) per ray.
int vxd_cpu_query(const cpu_scene* scene, const vx_camera* cam, const vx_ray* rays, uint32_t n, vx_hit* hits, int threads) {
  if (!scene || !cam || (n && (!rays || !hits))) return -1;
  const Scene S = to_scene(scene, nullptr);
  parallel_for(n, std::max(1, threads), [&](uint32_t i, int) {
    plain::Ctx C;
    C.S = S;
    C.cam = cam;
    C.grass_taper = nullptr;
    C.geom_only = true;
    plain::QueryRay q;
    const vx_ray& r = rays[i];
    plain::HitGeom h = plain::miss_hitgeom();
    if (plain::query_begin(C, F3(r.ox, r.oy, r.oz), F3(r.dx, r.dy, r.dz), r.tmax, q))
      while (!plain::query_step(C, q, h)) {
      }
    hits[i] = plain::query_result(h);
  });
  return 0;
}

}  // extern "C"
