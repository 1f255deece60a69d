// vx_runtime.hip — the renderer boundary of libvoxhip (C ABI in voxhip.h).
//
// Replaces render::Renderer (render/state/mod.rs:34-712): owns the resident
// world arenas and per-view history in HBM, applies region uploads, runs the
// pass sequence of Renderer::encode_compute (render/state/mod.rs:329-446) as
// stream-ordered kernel launches, and times each pass with HIP events.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <numeric>
#include <vector>

#include "vx_internal.h"
#include "vx_launch.h"

namespace {
thread_local char g_err[512] = "";
}

uint64_t vx_internal_node_capacity(const vx_ctx* ctx);
int vx_internal_device(const vx_ctx* ctx);

int vx_fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

#define VX_HIP(call)                                                                          \
  do {                                                                                        \
    hipError_t e_ = (call);                                                                   \
    if (e_ != hipSuccess) return vx_fail(VX_E_DEVICE, "%s: %s", #call, hipGetErrorString(e_)); \
  } while (0)

// Scratch of vx_render_views_noalloc, one copy per GPU for the life of the
// process: the fused launch's view table, the views' ping words and the
// per-band tile cursors, sized for VX_NOALLOC_MAX_VIEWS views.  Host code
// never uses the symbol itself as an address: vx_create resolves it once per
// device with hipGetSymbolAddress (with that device current) and the entry
// point passes the resolved pointer to the copies and kernels.
namespace {
struct NoAllocScratch {
  vxd::BatchView table[VX_NOALLOC_MAX_VIEWS];
  uint32_t ping[VX_NOALLOC_MAX_VIEWS];
  uint32_t cursor[2 * vxd::kBatchBands];
};
static __device__ __align__(256) unsigned char scratch[sizeof(NoAllocScratch)];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
NoAllocScratch* g_scratch_ptr[kMaxDevices] = {};

// Resolves this device's copy of `scratch` once (the caller has made `dev` current).
int resolve_scratch(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return vx_fail(VX_E_INVALID, "device %d out of range for the static scratch", dev);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[dev]) return VX_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) return vx_fail(VX_E_DEVICE, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
  g_scratch_ptr[dev] = static_cast<NoAllocScratch*>(p);
  return VX_OK;
}
}  // namespace

// The sort of a set of ray queries by their keys into grow-only device
// buffers: keys and indices, in and out (2 x cap each), and the radix sort's
// temporary storage.
struct RaySort {
  uint32_t* keys = nullptr;
  uint32_t* vals = nullptr;
  void* temp = nullptr;
  size_t temp_bytes = 0;
  uint32_t cap = 0;
};
void ray_sort_free(RaySort& r) {
  hipFree(r.keys);
  hipFree(r.vals);
  hipFree(r.temp);
  r = RaySort{};
}

struct vx_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  uint32_t flags = 0;
  uint64_t node_cap = 0;
  uint32_t chunk_cap = 0, grid_cap = 0, clip_levels = 0, clip_res = 0;
  vx_node* nodes = nullptr;
  vx_node_ropes* ropes = nullptr;
  vx_chunk_meta* meta = nullptr;
  uint32_t* macro = nullptr;
  uint32_t* colinfo = nullptr;
  uint32_t* grid = nullptr;
  float* clip_height = nullptr;
  vx_clipmap* clip_dev = nullptr;
  vx_clip_rect* clip_rects = nullptr;  // VX_CLIP_RECTS_MAX rectangles for the device clipmap fill
  // Per chunk slot: whether its nodes hold a LIGHT leaf (derived on the device
  // from the node arena and the chunk table after they change).
  uint32_t* chunk_light = nullptr;
  std::atomic<bool> lights_dirty{true};
  // The primary kernel's deferred grass tests: kDeferRegionsPerCU regions of
  // kDeferEntries x 64 entries per CU (32 KiB each; 128 MiB on 256 CUs).
  uint4* defer = nullptr;
  uint32_t defer_regions = 0;
  // Ray queries: the scene camera copy, the ray cursor, configuration, timing.
  vx_camera* qcam_dev = nullptr;
  uint32_t* qcursor = nullptr;
  unsigned long long* qcounters = nullptr;  // vx_query_count's counters (kViewCounters)
  RaySort qsort;  // the sorted order of a query call's rays
  vx_query_config qcfg{8u, 48u, 1u, 0u};
  hipEvent_t qev[2] = {nullptr, nullptr};
  // Live objects created on this context (views, batches, servers, chunk
  // builders).  vx_destroy refuses while any is alive: each of them launches
  // on the context's stream and reads or writes its arenas.
  std::atomic<int> children{0};
};

uint64_t vx_internal_node_capacity(const vx_ctx* ctx) { return ctx->node_cap; }
int vx_internal_device(const vx_ctx* ctx) { return ctx->device; }
void vx_internal_get_arenas(const vx_ctx* ctx, vx_internal_arenas* out) {
  out->nodes = ctx->nodes;
  out->ropes = ctx->ropes;
  out->macro = ctx->macro;
  out->colinfo = ctx->colinfo;
  out->node_cap = ctx->node_cap;
  out->chunk_cap = ctx->chunk_cap;
  out->stream = (void*)ctx->stream;
  out->device = ctx->device;
}
void vx_internal_nodes_changed(vx_ctx* ctx) { ctx->lights_dirty = true; }
void vx_internal_child(vx_ctx* ctx, int delta) { ctx->children.fetch_add(delta); }

struct vx_view {
  vx_ctx* ctx = nullptr;
  uint32_t pw = 0, ph = 0, rw = 0, rh = 0;
  int ping = 0;
  vx_camera* cam_dev = nullptr;
  float4* color = nullptr;
  float* depth = nullptr;
  float4* local = nullptr;
  float4* primary_hist[2] = {nullptr, nullptr};
  float* shadow[2] = {nullptr, nullptr};
  float4* local_hist[2] = {nullptr, nullptr};
  float4* godray[2] = {nullptr, nullptr};
  float4* pre_taa = nullptr;
  float4* output = nullptr;
  float4* out_hist[2] = {nullptr, nullptr};
  vx_hit* hits = nullptr;
  // One-view work-stealing launch: table [no hits, hits] and a small buffer
  // {0, 1, cursors[2 * 8]} whose first two words serve as the ping value.
  vxd::BatchView* ws_tab = nullptr;
  uint32_t* ws_buf = nullptr;
  unsigned long long* counters = nullptr;
  uint32_t* prof = nullptr;  // the five profile counters (ray_main.wgsl:32-47), kept across resizes
  uint32_t profile_flags = 0;  // of the camera last written (the launch picks the profiling kernel by it)
  float* ray_bytes = nullptr;  // per ray: counted bytes of the last counting primary pass (allocated on first use)
  // Slot written last by each history-producing pass.
  int written_primary = 1, written_local = 1, written_godray = 1, written_out = 1;
  hipEvent_t ev[10] = {};
  uint32_t last_mask = 0;
  uint32_t gen = 0;  // bumped when the view's buffers are reallocated
};

namespace {

// A view's counter words: the array counts, then the part byte totals.
constexpr size_t kViewCounters = 32;
// Ray query calls (and a batch's fused ray queries) of fewer rays trace them in
// input order: the sort's fixed cost (about 0.3 ms) outweighs the coherence it
// buys below about a million rays (measured: 512 Ki rays 1.68 ms unsorted
// against 1.86 sorted; 1 Mi rays 2.79 against 2.59; 2 M rays 5.19 against 4.02).
constexpr uint32_t kQuerySortMin = 1u << 20;

static_assert(vxd::kNumCounts + vxd::kNumParts <= kViewCounters, "view counters hold the counts and the parts");
static_assert(vxd::kNumParts == VX_COUNT_PARTS, "the ABI's part list is the kernels'");

struct DeviceGuard {
  int prev = 0;
  explicit DeviceGuard(int dev) {
    hipGetDevice(&prev);
    hipSetDevice(dev);
  }
  ~DeviceGuard() { hipSetDevice(prev); }
};

template <class T>
int dalloc(T** p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  hipError_t e = hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T));
  if (e != hipSuccess) return vx_fail(VX_E_NOMEM, "hipMalloc(%zu bytes): %s", count * sizeof(T), hipGetErrorString(e));
  e = hipMemset(*p, 0, count * sizeof(T));
  // The fill runs on the null stream, which the library's non-blocking
  // streams do not wait for: finish it before any of them touches the memory.
  if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
  if (e != hipSuccess) return vx_fail(VX_E_DEVICE, "hipMemset: %s", hipGetErrorString(e));
  return VX_OK;
}
template <class T>
void dfree(T*& p) {
  if (p) hipFree(p);
  p = nullptr;
}

vxd::Scene make_scene(const vx_ctx* c, bool godray) {
  vxd::Scene S;
  S.chunks = c->meta;
  S.nodes = reinterpret_cast<const uint4*>(c->nodes);
  S.ropes = reinterpret_cast<const uint4*>(c->ropes);
  S.grid = c->grid;
  S.macro = c->macro;
  S.colinfo = c->colinfo;
  S.clip = c->clip_dev;
  S.clip_height = c->clip_height;
  S.n_chunks = c->chunk_cap;
  S.n_nodes = c->node_cap;
  S.n_grid = c->grid_cap;
  S.n_macro = (uint64_t)c->chunk_cap * 16u;
  S.n_colinfo = (uint64_t)c->chunk_cap * 2048u;
  S.n_clip = (uint64_t)c->clip_levels * c->clip_res * c->clip_res;
  S.chunk_light = c->chunk_light;
  S.defer = c->defer;
  S.n_defer_regions = c->defer_regions;
  if (godray && !(c->flags & VX_CTX_GODRAY_MACRO_FIX)) {
    // The reference's godray pass sees the rope buffer where it expects macro_occ.
    S.macro = reinterpret_cast<const uint32_t*>(c->ropes);
    S.n_macro = c->node_cap * 8u;
  }
  return S;
}

// Recomputes the per-slot LIGHT flags on `st` if the node arena or the chunk
// table changed since they were last derived.
int refresh_chunk_lights(vx_ctx* c, hipStream_t st) {
  if (!c->lights_dirty.exchange(false)) return VX_OK;
  hipError_t e = vx_launch_chunk_lights(make_scene(c, false), c->chunk_light, c->chunk_cap, st);
  if (e != hipSuccess) {
    c->lights_dirty = true;
    return vx_fail(VX_E_DEVICE, "chunk light flags: %s", hipGetErrorString(e));
  }
  return VX_OK;
}

void free_view_buffers(vx_view* v) {
  dfree(v->color);
  dfree(v->depth);
  dfree(v->local);
  for (int i = 0; i < 2; ++i) {
    dfree(v->primary_hist[i]);
    dfree(v->shadow[i]);
    dfree(v->local_hist[i]);
    dfree(v->godray[i]);
    dfree(v->out_hist[i]);
  }
  dfree(v->pre_taa);
  dfree(v->output);
  dfree(v->hits);
  dfree(v->ws_tab);
  dfree(v->ws_buf);
  dfree(v->ray_bytes);
}

vxd::FrameBuffers view_buffers(const vx_view* v, int ping, bool hits);
vxd::BatchView one_view_entry(const vx_view* v, bool hits);

int alloc_view_buffers(vx_view* v) {
  const size_t r = (size_t)v->rw * v->rh, p = (size_t)v->pw * v->ph;
  int rc = VX_OK;
  if ((rc = dalloc(&v->color, r)) || (rc = dalloc(&v->depth, r)) || (rc = dalloc(&v->local, r))) return rc;
  for (int i = 0; i < 2; ++i) {
    if ((rc = dalloc(&v->primary_hist[i], r)) || (rc = dalloc(&v->shadow[i], r)) || (rc = dalloc(&v->local_hist[i], r)) ||
        (rc = dalloc(&v->godray[i], r)) || (rc = dalloc(&v->out_hist[i], p)))
      return rc;
  }
  if ((rc = dalloc(&v->pre_taa, p)) || (rc = dalloc(&v->output, p)) || (rc = dalloc(&v->hits, r))) return rc;
  if (!v->cam_dev) return vx_fail(VX_E_INVALID, "view tables need the camera buffer");
  if ((rc = dalloc(&v->ws_tab, 2)) || (rc = dalloc(&v->ws_buf, 2 + 2 * vxd::kBatchBands))) return rc;
  const vxd::BatchView tab[2] = {one_view_entry(v, false), one_view_entry(v, true)};
  const uint32_t pings[2] = {0u, 1u};
  if (hipMemcpy(v->ws_tab, tab, sizeof(tab), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(v->ws_buf, pings, sizeof(pings), hipMemcpyHostToDevice) != hipSuccess)
    return vx_fail(VX_E_DEVICE, "view table upload");
  v->ping = 0;
  v->written_primary = v->written_local = v->written_godray = v->written_out = 1;
  return VX_OK;
}

}  // namespace

extern "C" {

const char* vx_last_error(void) { return g_err; }

const char* vx_version(void) {
  return "libvoxhip 0.1 (hipcc " __VERSION__ ", gfx950, -O3 -ffp-contract=off, correctly rounded fp32 div/sqrt)";
}

int vx_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int vx_create(const vx_caps* caps, vx_ctx** out) {
  if (!out) return vx_fail(VX_E_INVALID, "vx_create: out is null");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return vx_fail(VX_E_NODEVICE, "vx_create: no GPU");
  vx_caps c{};
  if (caps) c = *caps;
  if (c.device < 0 || c.device >= ndev) return vx_fail(VX_E_INVALID, "vx_create: bad device %d", c.device);
  vx_ctx* ctx = new (std::nothrow) vx_ctx();
  if (!ctx) return vx_fail(VX_E_NOMEM, "vx_create: host allocation");
  ctx->device = c.device;
  if (c.flags & ~VX_CTX_GODRAY_MACRO_FIX) {
    delete ctx;
    return vx_fail(VX_E_INVALID, "vx_create: unknown flags 0x%x", c.flags);
  }
  ctx->flags = c.flags;
  ctx->node_cap = c.node_capacity ? c.node_capacity : (1024ull * 1024 * 1024) / sizeof(vx_node);  // buffers.rs:79
  ctx->chunk_cap = c.chunk_capacity ? c.chunk_capacity : 57u * 4u * 57u;                           // buffers.rs:87-96
  ctx->grid_cap = c.grid_capacity ? c.grid_capacity : 57u * 4u * 57u;
  ctx->clip_levels = c.clip_levels ? c.clip_levels : VX_CLIP_LEVELS_MAX;
  ctx->clip_res = c.clip_res ? c.clip_res : 256u;
  if (ctx->node_cap > 0xFFFFFFFFull || ctx->clip_levels > VX_CLIP_LEVELS_MAX) {
    delete ctx;
    return vx_fail(VX_E_INVALID, "vx_create: capacity out of range");
  }
  DeviceGuard g(ctx->device);
  int rc = VX_OK;
  if ((rc = dalloc(&ctx->nodes, ctx->node_cap)) || (rc = dalloc(&ctx->ropes, ctx->node_cap)) ||
      (rc = dalloc(&ctx->meta, ctx->chunk_cap)) || (rc = dalloc(&ctx->macro, (size_t)ctx->chunk_cap * 16)) ||
      (rc = dalloc(&ctx->colinfo, (size_t)ctx->chunk_cap * 2048)) || (rc = dalloc(&ctx->grid, ctx->grid_cap)) ||
      (rc = dalloc(&ctx->clip_height, (size_t)ctx->clip_levels * ctx->clip_res * ctx->clip_res)) ||
      (rc = dalloc(&ctx->clip_dev, 1)) || (rc = dalloc(&ctx->clip_rects, VX_CLIP_RECTS_MAX)) ||
      (rc = dalloc(&ctx->chunk_light, ctx->chunk_cap))) {
    vx_destroy(ctx);
    return rc;
  }
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device) != hipSuccess || cus <= 0) cus = 256;
  ctx->defer_regions = (uint32_t)cus * vxd::kDeferRegionsPerCU;
  if ((rc = dalloc(&ctx->defer, (size_t)ctx->defer_regions * vxd::kDeferEntries * 64u))) {
    vx_destroy(ctx);
    return rc;
  }
  vx_clipmap empty;
  std::memset(&empty, 0, sizeof(empty));
  empty.res = ctx->clip_res;
  if (hipMemcpy(ctx->clip_dev, &empty, sizeof(empty), hipMemcpyHostToDevice) != hipSuccess ||
      hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
    vx_destroy(ctx);
    return vx_fail(VX_E_DEVICE, "vx_create: stream/copy failed");
  }
  if (const int rc = resolve_scratch(ctx->device)) {
    vx_destroy(ctx);
    return rc;
  }
  *out = ctx;
  return VX_OK;
}

int vx_destroy(vx_ctx* ctx) {
  if (!ctx) return VX_OK;
  const int live = ctx->children.load();
  if (live > 0)
    return vx_fail(VX_E_BUSY, "vx_destroy: %d view/batch/server/chunk-builder object(s) of this context are still alive; "
                   "destroy them first", live);
  DeviceGuard g(ctx->device);
  if (ctx->stream) {
    hipStreamSynchronize(ctx->stream);
    hipStreamDestroy(ctx->stream);
  }
  dfree(ctx->nodes);
  dfree(ctx->ropes);
  dfree(ctx->meta);
  dfree(ctx->macro);
  dfree(ctx->colinfo);
  dfree(ctx->grid);
  dfree(ctx->clip_height);
  dfree(ctx->clip_dev);
  dfree(ctx->clip_rects);
  dfree(ctx->chunk_light);
  dfree(ctx->defer);
  dfree(ctx->qcam_dev);
  dfree(ctx->qcursor);
  dfree(ctx->qcounters);
  ray_sort_free(ctx->qsort);
  for (auto& e : ctx->qev)
    if (e) hipEventDestroy(e);
  delete ctx;
  return VX_OK;
}

int vx_upload_regions(vx_ctx* ctx, int arena, const vx_region* regions, uint32_t n) {
  if (!ctx || (!regions && n)) return vx_fail(VX_E_INVALID, "vx_upload_regions: null argument");
  char* base = nullptr;
  size_t esz = 0;
  uint64_t cap = 0;
  switch (arena) {
    case VX_ARENA_NODES: base = (char*)ctx->nodes; esz = sizeof(vx_node); cap = ctx->node_cap; break;
    case VX_ARENA_ROPES: base = (char*)ctx->ropes; esz = sizeof(vx_node_ropes); cap = ctx->node_cap; break;
    case VX_ARENA_META: base = (char*)ctx->meta; esz = sizeof(vx_chunk_meta); cap = ctx->chunk_cap; break;
    case VX_ARENA_MACRO: base = (char*)ctx->macro; esz = 4; cap = (uint64_t)ctx->chunk_cap * 16; break;
    case VX_ARENA_COLINFO: base = (char*)ctx->colinfo; esz = 4; cap = (uint64_t)ctx->chunk_cap * 2048; break;
    default: return vx_fail(VX_E_INVALID, "vx_upload_regions: bad arena %d", arena);
  }
  for (uint32_t i = 0; i < n; ++i) {
    const vx_region& r = regions[i];
    if (r.count == 0) continue;
    if (!r.data) return vx_fail(VX_E_INVALID, "vx_upload_regions: region %u has no data", i);
    if (r.offset > cap || r.count > cap - r.offset)
      return vx_fail(VX_E_CAPACITY, "vx_upload_regions: region %u [%llu, +%llu) exceeds capacity %llu", i,
                     (unsigned long long)r.offset, (unsigned long long)r.count, (unsigned long long)cap);
  }
  DeviceGuard g(ctx->device);
  for (uint32_t i = 0; i < n; ++i) {
    const vx_region& r = regions[i];
    if (r.count == 0) continue;
    VX_HIP(hipMemcpyAsync(base + r.offset * esz, r.data, r.count * esz, hipMemcpyHostToDevice, ctx->stream));
  }
  VX_HIP(hipStreamSynchronize(ctx->stream));
  if (arena == VX_ARENA_NODES || arena == VX_ARENA_META) ctx->lights_dirty = true;
  return VX_OK;
}

int vx_read_region(vx_ctx* ctx, int arena, uint64_t offset, uint64_t count, void* dst) {
  if (!ctx || (!dst && count)) return vx_fail(VX_E_INVALID, "vx_read_region: null argument");
  const char* base = nullptr;
  size_t esz = 0;
  uint64_t cap = 0;
  switch (arena) {
    case VX_ARENA_NODES: base = (const char*)ctx->nodes; esz = sizeof(vx_node); cap = ctx->node_cap; break;
    case VX_ARENA_ROPES: base = (const char*)ctx->ropes; esz = sizeof(vx_node_ropes); cap = ctx->node_cap; break;
    case VX_ARENA_META: base = (const char*)ctx->meta; esz = sizeof(vx_chunk_meta); cap = ctx->chunk_cap; break;
    case VX_ARENA_MACRO: base = (const char*)ctx->macro; esz = 4; cap = (uint64_t)ctx->chunk_cap * 16; break;
    case VX_ARENA_COLINFO: base = (const char*)ctx->colinfo; esz = 4; cap = (uint64_t)ctx->chunk_cap * 2048; break;
    default: return vx_fail(VX_E_INVALID, "vx_read_region: bad arena %d", arena);
  }
  if (offset > cap || count > cap - offset)
    return vx_fail(VX_E_CAPACITY, "vx_read_region: [%llu, +%llu) exceeds capacity %llu", (unsigned long long)offset,
                   (unsigned long long)count, (unsigned long long)cap);
  if (!count) return VX_OK;
  DeviceGuard g(ctx->device);
  VX_HIP(hipMemcpyAsync(dst, base + offset * esz, count * esz, hipMemcpyDeviceToHost, ctx->stream));
  VX_HIP(hipStreamSynchronize(ctx->stream));
  return VX_OK;
}

int vx_write_grid(vx_ctx* ctx, const uint32_t* grid, uint32_t cells) {
  if (!ctx || (!grid && cells)) return vx_fail(VX_E_INVALID, "vx_write_grid: null argument");
  if (cells > ctx->grid_cap) return vx_fail(VX_E_CAPACITY, "vx_write_grid: %u cells > capacity %u", cells, ctx->grid_cap);
  DeviceGuard g(ctx->device);
  VX_HIP(hipMemcpyAsync(ctx->grid, grid, (size_t)cells * 4, hipMemcpyHostToDevice, ctx->stream));
  VX_HIP(hipStreamSynchronize(ctx->stream));
  return VX_OK;
}

int vx_write_clipmap(vx_ctx* ctx, const vx_clipmap* params, const vx_clip_patch* patches, uint32_t n) {
  if (!ctx || !params || (!patches && n)) return vx_fail(VX_E_INVALID, "vx_write_clipmap: null argument");
  if (params->levels > ctx->clip_levels || (params->levels && params->res != ctx->clip_res))
    return vx_fail(VX_E_CAPACITY, "vx_write_clipmap: levels/res (%u, %u) do not fit (%u, %u)", params->levels, params->res,
                   ctx->clip_levels, ctx->clip_res);
  for (uint32_t i = 0; i < n; ++i) {
    const vx_clip_patch& p = patches[i];
    if (p.level >= ctx->clip_levels || p.x > ctx->clip_res || p.w > ctx->clip_res - p.x || p.y > ctx->clip_res ||
        p.h > ctx->clip_res - p.y || (!p.data && p.w && p.h))
      return vx_fail(VX_E_INVALID, "vx_write_clipmap: patch %u out of range", i);
  }
  DeviceGuard g(ctx->device);
  for (uint32_t i = 0; i < n; ++i) {
    const vx_clip_patch& p = patches[i];
    if (!p.w || !p.h) continue;
    float* dst = ctx->clip_height + ((size_t)p.level * ctx->clip_res + p.y) * ctx->clip_res + p.x;
    VX_HIP(hipMemcpy2DAsync(dst, (size_t)ctx->clip_res * 4, p.data, (size_t)p.w * 4, (size_t)p.w * 4, p.h, hipMemcpyHostToDevice,
                            ctx->stream));
  }
  VX_HIP(hipMemcpyAsync(ctx->clip_dev, params, sizeof(vx_clipmap), hipMemcpyHostToDevice, ctx->stream));
  VX_HIP(hipStreamSynchronize(ctx->stream));
  return VX_OK;
}

int vx_write_clipmap_rects(vx_ctx* ctx, const vx_clipmap* params, const vx_world_params* world, const vx_clip_rect* rects,
                           uint32_t n) {
  if (!ctx || !params || !world || (!rects && n)) return vx_fail(VX_E_INVALID, "vx_write_clipmap_rects: null argument");
  if (params->levels > ctx->clip_levels || (params->levels && params->res != ctx->clip_res))
    return vx_fail(VX_E_CAPACITY, "vx_write_clipmap_rects: levels/res (%u, %u) do not fit (%u, %u)", params->levels,
                   params->res, ctx->clip_levels, ctx->clip_res);
  if (n > VX_CLIP_RECTS_MAX)
    return vx_fail(VX_E_INVALID, "vx_write_clipmap_rects: %u rectangles (at most %u)", n, VX_CLIP_RECTS_MAX);
  uint32_t max_texels = 0;
  for (uint32_t i = 0; i < n; ++i) {
    const vx_clip_rect& r = rects[i];
    if (r.level >= ctx->clip_levels || r.res != ctx->clip_res || r.w > r.res || r.h > r.res || r.x >= r.res ||
        r.y >= r.res || r.x + r.w > r.res || r.y + r.h > r.res || r.lx0 >= r.res || r.lz0 >= r.res)
      return vx_fail(VX_E_INVALID, "vx_write_clipmap_rects: rectangle %u out of range", i);
    max_texels = std::max(max_texels, r.w * r.h);
  }
  DeviceGuard g(ctx->device);
  if (n && max_texels) {
    VX_HIP(hipMemcpyAsync(ctx->clip_rects, rects, n * sizeof(vx_clip_rect), hipMemcpyHostToDevice, ctx->stream));
    VX_HIP(vx_launch_clip_fill(ctx->clip_rects, n, max_texels, *world, ctx->clip_height, ctx->stream));
  }
  VX_HIP(hipMemcpyAsync(ctx->clip_dev, params, sizeof(vx_clipmap), hipMemcpyHostToDevice, ctx->stream));
  VX_HIP(hipStreamSynchronize(ctx->stream));
  return VX_OK;
}

int vx_read_clipmap(vx_ctx* ctx, float* heights, size_t bytes) {
  if (!ctx || !heights) return vx_fail(VX_E_INVALID, "vx_read_clipmap: null argument");
  const size_t need = (size_t)ctx->clip_levels * ctx->clip_res * ctx->clip_res * sizeof(float);
  if (bytes < need) return vx_fail(VX_E_INVALID, "vx_read_clipmap: %zu bytes, need %zu", bytes, need);
  DeviceGuard g(ctx->device);
  VX_HIP(hipMemcpyAsync(heights, ctx->clip_height, need, hipMemcpyDeviceToHost, ctx->stream));
  VX_HIP(hipStreamSynchronize(ctx->stream));
  return VX_OK;
}

int vx_view_create(vx_ctx* ctx, uint32_t pw, uint32_t ph, uint32_t rw, uint32_t rh, vx_view** out) {
  if (!ctx || !out || !pw || !ph || !rw || !rh) return vx_fail(VX_E_INVALID, "vx_view_create: bad argument");
  if ((uint64_t)pw * ph > (1ull << 28) || (uint64_t)rw * rh > (1ull << 28)) return vx_fail(VX_E_INVALID, "vx_view_create: too large");
  *out = nullptr;
  vx_view* v = new (std::nothrow) vx_view();
  if (!v) return vx_fail(VX_E_NOMEM, "vx_view_create: host allocation");
  v->ctx = ctx;
  ctx->children.fetch_add(1);  // released by vx_view_destroy, also on the failure path below
  v->pw = pw;
  v->ph = ph;
  v->rw = rw;
  v->rh = rh;
  DeviceGuard g(ctx->device);
  // The camera buffer and the profile counters first: the view's kernel table
  // (built with the buffers) points at them.
  int rc = dalloc(&v->cam_dev, 1);
  if (!rc) rc = dalloc(&v->prof, 5);
  if (!rc) rc = alloc_view_buffers(v);
  if (!rc) rc = dalloc(&v->counters, kViewCounters);
  for (int i = 0; i < 10 && !rc; ++i)
    if (hipEventCreate(&v->ev[i]) != hipSuccess) rc = vx_fail(VX_E_DEVICE, "vx_view_create: hipEventCreate");
  if (rc) {
    vx_view_destroy(v);
    return rc;
  }
  *out = v;
  return VX_OK;
}

void vx_view_destroy(vx_view* v) {
  if (!v) return;
  DeviceGuard g(v->ctx->device);
  hipStreamSynchronize(v->ctx->stream);
  free_view_buffers(v);
  dfree(v->cam_dev);
  dfree(v->prof);
  dfree(v->counters);
  for (int i = 0; i < 10; ++i)
    if (v->ev[i]) hipEventDestroy(v->ev[i]);
  v->ctx->children.fetch_sub(1);
  delete v;
}

int vx_view_resize(vx_view* v, uint32_t pw, uint32_t ph, uint32_t rw, uint32_t rh) {
  if (!v || !pw || !ph || !rw || !rh) return vx_fail(VX_E_INVALID, "vx_view_resize: bad argument");
  if ((uint64_t)pw * ph > (1ull << 28) || (uint64_t)rw * rh > (1ull << 28)) return vx_fail(VX_E_INVALID, "vx_view_resize: too large");
  DeviceGuard g(v->ctx->device);
  VX_HIP(hipStreamSynchronize(v->ctx->stream));
  free_view_buffers(v);
  v->pw = pw;
  v->ph = ph;
  v->rw = rw;
  v->rh = rh;
  ++v->gen;
  return alloc_view_buffers(v);
}

int vx_view_reset(vx_view* v) {
  if (!v) return vx_fail(VX_E_INVALID, "vx_view_reset: null view");
  DeviceGuard g(v->ctx->device);
  const size_t r = (size_t)v->rw * v->rh, p = (size_t)v->pw * v->ph;
  hipStream_t s = v->ctx->stream;
  for (int i = 0; i < 2; ++i) {
    VX_HIP(hipMemsetAsync(v->primary_hist[i], 0, r * 16, s));
    VX_HIP(hipMemsetAsync(v->shadow[i], 0, r * 4, s));
    VX_HIP(hipMemsetAsync(v->local_hist[i], 0, r * 16, s));
    VX_HIP(hipMemsetAsync(v->godray[i], 0, r * 16, s));
    VX_HIP(hipMemsetAsync(v->out_hist[i], 0, p * 16, s));
  }
  VX_HIP(hipStreamSynchronize(s));
  v->ping = 0;
  v->written_primary = v->written_local = v->written_godray = v->written_out = 1;
  return VX_OK;
}

int vx_write_camera(vx_view* v, const vx_camera* cam) {
  if (!v || !cam) return vx_fail(VX_E_INVALID, "vx_write_camera: null argument");
  DeviceGuard g(v->ctx->device);
  VX_HIP(hipMemcpyAsync(v->cam_dev, cam, sizeof(vx_camera), hipMemcpyHostToDevice, v->ctx->stream));
  VX_HIP(hipStreamSynchronize(v->ctx->stream));
  v->profile_flags = cam->profile_flags;
  return VX_OK;
}

}  // extern "C"

struct vx_batch {
  vx_ctx* ctx = nullptr;
  std::vector<vx_view*> views;      // table order (largest render first)
  std::vector<uint32_t> gens;       // view buffer generations at creation
  std::vector<uint32_t> ping_host;  // mirror of the device ping array
  vxd::BatchView* table = nullptr;
  uint32_t* ping = nullptr;
  uint32_t* cursor = nullptr;
  vxd::Batch B{};
  vx_batch_config cfg{6u, 1u, 1u, 0u};
  std::map<uint64_t, hipGraphExec_t> graphs;  // by (mask, config)
  hipEvent_t ev[2] = {nullptr, nullptr};
  uint32_t cap = 0;                      // views the device table and ping arrays hold
  std::vector<vxd::BatchView> tab_host;  // the table as last uploaded
  // Ray queries traced by the fused primary launch (vx_batch_set_rays).
  vxd::BatchRays* rays_dev = nullptr;
  vxd::BatchRays* rays_pin = nullptr;  // pinned host copy, uploaded before a launch when it changed
  hipEvent_t rays_ev = nullptr;        // recorded after the last upload read rays_pin
  bool rays_dirty = false;
  uint32_t rays_n = 0;
  bool rays_sort = false;  // each render sorts the rays first (RaySort below)
  RaySort rsort;
  // Candidate lists of every 8x8 tile for the execution tile shapes other
  // than 8x8 (allocated on the first such render, grown with the views).
  uint32_t* lists = nullptr;
  uint32_t lists_cap = 0;  // tiles the list buffer holds
};

namespace {
uint32_t persistent_groups(const vx_batch_config& cfg) {
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) == hipSuccess) {
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && n > 0) cus = n;
  }
  return (uint32_t)cus * 4u * std::max(1u, cfg.waves_per_simd);
}
void drop_graphs(vx_batch* b) {
  for (auto& kv : b->graphs) hipGraphExecDestroy(kv.second);
  b->graphs.clear();
}
int check_batch_views(const vx_ctx* ctx, vx_view* const* views, uint32_t n, const char* fn) {
  for (uint32_t i = 0; i < n; ++i) {
    if (!views[i] || views[i]->ctx != ctx) return vx_fail(VX_E_INVALID, "%s: view %u is null or of another context", fn, i);
    for (uint32_t j = 0; j < i; ++j)
      if (views[j] == views[i]) return vx_fail(VX_E_INVALID, "%s: view %u appears twice", fn, i);
  }
  return VX_OK;
}
}  // namespace

namespace {

vxd::FrameBuffers view_buffers(const vx_view* v, int ping, bool hits) {
  const int pong = 1 - ping;
  vxd::FrameBuffers fb;
  std::memset(&fb, 0, sizeof(fb));
  fb.rw = v->rw;
  fb.rh = v->rh;
  fb.pw = v->pw;
  fb.ph = v->ph;
  fb.primary_hist_in = v->primary_hist[ping];
  fb.primary_hist_out = v->primary_hist[pong];
  fb.shadow_in = v->shadow[ping];
  fb.shadow_out = v->shadow[pong];
  fb.color = v->color;
  fb.depth = v->depth;
  fb.local = v->local;
  fb.local_hist_in = v->local_hist[ping];
  fb.local_hist_out = v->local_hist[pong];
  fb.godray_hist = v->godray[ping];
  fb.godray_out = v->godray[pong];
  fb.pre_taa = v->pre_taa;
  fb.out_hist_in = v->out_hist[ping];
  fb.out_hist_out = v->out_hist[pong];
  fb.output = v->output;
  fb.hits = hits ? v->hits : nullptr;
  fb.prof = v->prof;
  return fb;
}

// Binds a batch to `views`: the view table (largest render first: their tiles
// are taken first, small views fill the tail) and the ping mirror, uploaded
// before this returns.  The device arrays grow only when n exceeds their
// capacity; graphs captured for the old table are dropped.
// Table entries of a fused launch over `views` (in table order) and the
// launch totals: tiles, render blocks, present blocks.
// The views' table entries and their work item totals (8x8 tiles, render
// and present blocks); xtiles, if given, receives the execution tiles of
// shapes 16x4 and 32x2.
void batch_entries(vx_view* const* views, uint32_t n, vxd::BatchView* tab, uint32_t& tiles, uint32_t& rblks,
                   uint32_t& pblks, uint32_t* xtiles = nullptr) {
  tiles = rblks = pblks = 0;
  uint32_t xt[2] = {0u, 0u};
  for (uint32_t i = 0; i < n; ++i) {
    const vx_view* v = views[i];
    vxd::BatchView& e = tab[i];
    std::memset(&e, 0, sizeof(e));
    e.fb[0] = view_buffers(v, 0, false);
    e.fb[1] = view_buffers(v, 1, false);
    e.cam = v->cam_dev;
    e.tiles_x = (v->rw + 7u) / 8u;
    e.tiles = e.tiles_x * ((v->rh + 7u) / 8u);
    e.rblk_x = (v->rw + 15u) / 16u;
    e.rblks = e.rblk_x * ((v->rh + 15u) / 16u);
    e.pblk_x = (v->pw + 15u) / 16u;
    e.pblks = e.pblk_x * ((v->ph + 15u) / 16u);
    e.tile_base = tiles;
    e.rblk_base = rblks;
    e.pblk_base = pblks;
    tiles += e.tiles;
    rblks += e.rblks;
    pblks += e.pblks;
    for (uint32_t k = 0; k < 2u; ++k) {  // 16x4, 32x2
      const uint32_t tw = 16u << k, th = 64u / tw;
      e.xt_x[k] = (v->rw + tw - 1u) / tw;
      e.xt_base[k] = xt[k];
      xt[k] += e.xt_x[k] * ((v->rh + th - 1u) / th);
    }
  }
  if (xtiles) xtiles[0] = xt[0], xtiles[1] = xt[1];
}

bool larger_render(const vx_view* a, const vx_view* c) { return (uint64_t)a->rw * a->rh > (uint64_t)c->rw * c->rh; }

int batch_bind(vx_batch* b, vx_view* const* views, uint32_t n) {
  drop_graphs(b);
  b->views.assign(views, views + n);
  std::stable_sort(b->views.begin(), b->views.end(), larger_render);
  b->tab_host.resize(n);
  b->gens.clear();
  b->ping_host.clear();
  uint32_t tiles = 0, rblks = 0, pblks = 0, xtiles[2] = {0u, 0u};
  batch_entries(b->views.data(), n, b->tab_host.data(), tiles, rblks, pblks, xtiles);
  for (const vx_view* v : b->views) {
    b->gens.push_back(v->gen);
    b->ping_host.push_back((uint32_t)v->ping);
  }
  if (n > b->cap) {
    dfree(b->table);
    dfree(b->ping);
    b->table = nullptr;
    b->ping = nullptr;
    b->cap = 0;
    int rc = dalloc(&b->table, n);
    if (!rc) rc = dalloc(&b->ping, n);
    if (rc) return rc;
    b->cap = n;
  }
  hipStream_t st = b->ctx->stream;
  VX_HIP(hipMemcpyAsync(b->table, b->tab_host.data(), n * sizeof(vxd::BatchView), hipMemcpyHostToDevice, st));
  VX_HIP(hipMemcpyAsync(b->ping, b->ping_host.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, st));
  VX_HIP(hipStreamSynchronize(st));  // a render may run on another stream
  b->B.views = b->table;
  b->B.ping = b->ping;
  b->B.cursor = b->cursor;
  b->B.n = n;
  b->B.tiles = tiles;
  b->B.rblks = rblks;
  b->B.pblks = pblks;
  b->B.grab = b->cfg.grab;
  b->B.xtiles[0] = xtiles[0];
  b->B.xtiles[1] = xtiles[1];
  return VX_OK;
}

vxd::BatchView one_view_entry(const vx_view* v, bool hits) {
  vxd::BatchView e;
  std::memset(&e, 0, sizeof(e));
  e.fb[0] = view_buffers(v, 0, hits);
  e.fb[1] = view_buffers(v, 1, hits);
  e.cam = v->cam_dev;
  e.tiles_x = (v->rw + 7u) / 8u;
  e.tiles = e.tiles_x * ((v->rh + 7u) / 8u);
  e.rblk_x = (v->rw + 15u) / 16u;
  e.rblks = e.rblk_x * ((v->rh + 15u) / 16u);
  e.pblk_x = (v->pw + 15u) / 16u;
  e.pblks = e.pblk_x * ((v->ph + 15u) / 16u);
  return e;
}

}  // namespace

extern "C" {

int vx_render(vx_view* v, uint32_t mask, void* stream) {
  if (!v) return vx_fail(VX_E_INVALID, "vx_render: null view");
  vx_ctx* c = v->ctx;
  hipStream_t st = stream ? (hipStream_t)stream : c->stream;
  DeviceGuard g(c->device);
  const int ping = v->ping, pong = 1 - ping;
  const vxd::FrameBuffers fb = view_buffers(v, ping, (mask & VX_PASS_WRITE_HITS) != 0);
  const bool count = (mask & VX_PASS_COUNT) != 0;
  const vxd::Scene S = make_scene(c, false);
  // The primary pass runs as a persistent work-stealing kernel over the view's
  // tiles (the one-view case of a batch); the counting form is one workgroup
  // per tile.
  vxd::Batch B1{};
  B1.views = v->ws_tab + ((mask & VX_PASS_WRITE_HITS) ? 1 : 0);
  B1.ping = v->ws_buf + ping;  // ws_buf[0] = 0, ws_buf[1] = 1
  B1.cursor = v->ws_buf + 2;
  B1.n = 1;
  B1.tiles = ((v->rw + 7u) / 8u) * ((v->rh + 7u) / 8u);
  B1.grab = 1;
  // never more persistent workgroups than tiles
  const uint32_t persistent = std::min(persistent_groups(vx_batch_config{8u, 1u, 0u, 0u}), std::max(B1.tiles, 1u));
  if (mask & VX_PASS_PRIMARY) {
    const int rc = refresh_chunk_lights(c, st);
    if (rc) return rc;
  }
  VX_HIP(hipEventRecord(v->ev[0], st));
  if (!count && (mask & VX_PASS_PRIMARY)) VX_HIP(vx_launch_ws_begin(B1, st));
  if (mask & VX_PASS_PRIMARY) {
    if (count) {
      if (!v->ray_bytes) {
        const int rc = dalloc(&v->ray_bytes, (size_t)v->rw * v->rh);
        if (rc) return rc;
      }
      VX_HIP(vx_launch_primary_count(S, v->cam_dev, fb, (mask & VX_PASS_COUNT_EXECUTED) != 0u, v->counters,
                                     v->ray_bytes, st));
    } else {
      VX_HIP(vx_launch_primary_ws(S, B1, persistent, (v->profile_flags & 1u) != 0u, st));
    }
    v->written_primary = pong;
  }
  VX_HIP(hipEventRecord(v->ev[1], st));
  if (mask & VX_PASS_LOCAL_TAA) {
    VX_HIP(vx_launch_local_taa(fb, st));
    v->written_local = pong;
  }
  VX_HIP(hipEventRecord(v->ev[2], st));
  if (mask & VX_PASS_GODRAY) {
    VX_HIP(vx_launch_godray(make_scene(c, true), v->cam_dev, fb, count, v->counters, st));
    v->written_godray = pong;
  }
  VX_HIP(hipEventRecord(v->ev[3], st));
  if (mask & VX_PASS_COMPOSITE) VX_HIP(vx_launch_composite(S, v->cam_dev, fb, st));
  VX_HIP(hipEventRecord(v->ev[4], st));
  if (mask & VX_PASS_COMPOSITE_TAA) {
    VX_HIP(vx_launch_composite_taa(S, v->cam_dev, fb, st));
    v->written_out = pong;
  }
  VX_HIP(hipEventRecord(v->ev[5], st));
  v->last_mask = mask;
  if (mask & VX_PASS_END_FRAME) v->ping = pong;
  return VX_OK;
}

static int output_buffer(vx_view* v, int which, void** p, size_t* bytes) {
  const size_t r = (size_t)v->rw * v->rh, pp = (size_t)v->pw * v->ph;
  switch (which) {
    case VX_OUT_COLOR: *p = v->color; *bytes = r * 16; return VX_OK;
    case VX_OUT_DEPTH: *p = v->depth; *bytes = r * 4; return VX_OK;
    case VX_OUT_LOCAL: *p = v->local; *bytes = r * 16; return VX_OK;
    case VX_OUT_PRIMARY_HIST: *p = v->primary_hist[v->written_primary]; *bytes = r * 16; return VX_OK;
    case VX_OUT_SHADOW_HIST: *p = v->shadow[v->written_primary]; *bytes = r * 4; return VX_OK;
    case VX_OUT_LOCAL_HIST: *p = v->local_hist[v->written_local]; *bytes = r * 16; return VX_OK;
    case VX_OUT_GODRAY: *p = v->godray[v->written_godray]; *bytes = r * 16; return VX_OK;
    case VX_OUT_PRE_TAA: *p = v->pre_taa; *bytes = pp * 16; return VX_OK;
    case VX_OUT_OUTPUT: *p = v->output; *bytes = pp * 16; return VX_OK;
    case VX_OUT_HITS: *p = v->hits; *bytes = r * sizeof(vx_hit); return VX_OK;
    case VX_OUT_RAY_BYTES:
      if (!v->ray_bytes) return vx_fail(VX_E_INVALID, "VX_OUT_RAY_BYTES: no counting primary pass has run on this view");
      *p = v->ray_bytes;
      *bytes = r * 4;
      return VX_OK;
    default: return vx_fail(VX_E_INVALID, "bad output selector %d", which);
  }
}

int vx_read_output(vx_view* v, int which, void* dst, size_t bytes) {
  if (!v || !dst) return vx_fail(VX_E_INVALID, "vx_read_output: null argument");
  void* p = nullptr;
  size_t n = 0;
  int rc = output_buffer(v, which, &p, &n);
  if (rc) return rc;
  if (bytes < n) return vx_fail(VX_E_INVALID, "vx_read_output: need %zu bytes, got %zu", n, bytes);
  DeviceGuard g(v->ctx->device);
  VX_HIP(hipStreamSynchronize(v->ctx->stream));
  VX_HIP(hipDeviceSynchronize());
  VX_HIP(hipMemcpy(dst, p, n, hipMemcpyDeviceToHost));
  return VX_OK;
}

int vx_output_device_ptr(vx_view* v, int which, void** dptr, size_t* bytes) {
  if (!v || !dptr || !bytes) return vx_fail(VX_E_INVALID, "vx_output_device_ptr: null argument");
  return output_buffer(v, which, dptr, bytes);
}

int vx_write_history(vx_view* v, int which, const void* src, size_t bytes) {
  if (!v || !src) return vx_fail(VX_E_INVALID, "vx_write_history: null argument");
  const size_t r = (size_t)v->rw * v->rh, pp = (size_t)v->pw * v->ph;
  void* dst = nullptr;
  size_t n = 0;
  const int ping = v->ping;  // the slot the next frame reads
  switch (which) {
    case VX_OUT_PRIMARY_HIST: dst = v->primary_hist[ping]; n = r * 16; break;
    case VX_OUT_SHADOW_HIST: dst = v->shadow[ping]; n = r * 4; break;
    case VX_OUT_LOCAL_HIST: dst = v->local_hist[ping]; n = r * 16; break;
    case VX_OUT_GODRAY: dst = v->godray[ping]; n = r * 16; break;
    case VX_OUT_OUTPUT: dst = v->out_hist[ping]; n = pp * 16; break;
    default: return vx_fail(VX_E_INVALID, "vx_write_history: selector %d is not a history", which);
  }
  if (bytes != n) return vx_fail(VX_E_INVALID, "vx_write_history: need %zu bytes, got %zu", n, bytes);
  DeviceGuard g(v->ctx->device);
  VX_HIP(hipMemcpy(dst, src, n, hipMemcpyHostToDevice));
  return VX_OK;
}

int vx_view_counters(vx_view* v, uint64_t out[9], int reset) {
  if (!v || !out) return vx_fail(VX_E_INVALID, "vx_view_counters: null argument");
  DeviceGuard g(v->ctx->device);
  VX_HIP(hipDeviceSynchronize());
  unsigned long long tmp[kViewCounters];
  VX_HIP(hipMemcpy(tmp, v->counters, sizeof(tmp), hipMemcpyDeviceToHost));
  for (int i = 0; i < 9; ++i) out[i] = tmp[i];
  if (reset) {
    VX_HIP(hipMemset(v->counters, 0, sizeof(tmp)));
    VX_HIP(hipStreamSynchronize(nullptr));  // before the view's non-blocking stream counts again
  }
  return VX_OK;
}

int vx_view_count_parts(vx_view* v, uint64_t out[VX_COUNT_PARTS]) {
  if (!v || !out) return vx_fail(VX_E_INVALID, "vx_view_count_parts: null argument");
  DeviceGuard g(v->ctx->device);
  VX_HIP(hipDeviceSynchronize());
  unsigned long long tmp[kViewCounters];
  VX_HIP(hipMemcpy(tmp, v->counters, sizeof(tmp), hipMemcpyDeviceToHost));
  for (int i = 0; i < VX_COUNT_PARTS; ++i) out[i] = tmp[vxd::kNumCounts + i];
  return VX_OK;
}

int vx_view_profile(vx_view* v, uint32_t out[5], int reset) {
  if (!v || !out) return vx_fail(VX_E_INVALID, "vx_view_profile: null argument");
  DeviceGuard g(v->ctx->device);
  VX_HIP(hipDeviceSynchronize());  // single-view and batch renders, on any stream
  VX_HIP(hipMemcpy(out, v->prof, 5 * sizeof(uint32_t), hipMemcpyDeviceToHost));
  if (reset) {
    const uint32_t zero[5] = {0u, 0u, 0u, 0u, 0u};
    VX_HIP(hipMemcpy(v->prof, zero, sizeof(zero), hipMemcpyHostToDevice));
  }
  return VX_OK;
}

int vx_view_timings(vx_view* v, float ms[5]) {
  if (!v || !ms) return vx_fail(VX_E_INVALID, "vx_view_timings: null argument");
  DeviceGuard g(v->ctx->device);
  VX_HIP(hipEventSynchronize(v->ev[5]));
  for (int i = 0; i < 5; ++i) {
    float t = 0.0f;
    VX_HIP(hipEventElapsedTime(&t, v->ev[i], v->ev[i + 1]));
    ms[i] = t;
  }
  return VX_OK;
}

// ---------------------------------------------------------------------------
// Batches: many views rendered by fused launches (one kernel per pass over
// all views), optionally replayed from a captured hipGraph.
// ---------------------------------------------------------------------------


int vx_batch_create(vx_ctx* ctx, vx_view* const* views, uint32_t n, vx_batch** out) {
  if (!ctx || !views || !n || !out) return vx_fail(VX_E_INVALID, "vx_batch_create: bad argument");
  *out = nullptr;
  int rc = check_batch_views(ctx, views, n, "vx_batch_create");
  if (rc) return rc;
  DeviceGuard g(ctx->device);
  vx_batch* b = new (std::nothrow) vx_batch();
  if (!b) return vx_fail(VX_E_NOMEM, "vx_batch_create: host allocation");
  b->ctx = ctx;
  ctx->children.fetch_add(1);  // released by vx_batch_destroy
  rc = dalloc(&b->cursor, 2u * vxd::kBatchBands);
  for (int i = 0; i < 2 && !rc; ++i)
    if (hipEventCreate(&b->ev[i]) != hipSuccess) rc = vx_fail(VX_E_DEVICE, "vx_batch_create: hipEventCreate");
  if (!rc) rc = batch_bind(b, views, n);
  if (rc) {
    vx_batch_destroy(b);
    return rc;
  }
  *out = b;
  return VX_OK;
}

void vx_batch_destroy(vx_batch* b) {
  if (!b) return;
  DeviceGuard g(b->ctx->device);
  hipStreamSynchronize(b->ctx->stream);
  drop_graphs(b);
  dfree(b->table);
  dfree(b->ping);
  dfree(b->cursor);
  dfree(b->rays_dev);
  ray_sort_free(b->rsort);
  dfree(b->lists);
  if (b->rays_pin) hipHostFree(b->rays_pin);
  if (b->rays_ev) hipEventDestroy(b->rays_ev);
  for (auto& e : b->ev)
    if (e) hipEventDestroy(e);
  b->ctx->children.fetch_sub(1);
  delete b;
}

// Buffers for n rays (waits for `st` before replacing smaller ones, which a
// launch on it may still read).
int ray_sort_reserve(RaySort& r, uint32_t n, hipStream_t st) {
  if (n <= r.cap) return VX_OK;
  VX_HIP(hipStreamSynchronize(st));
  ray_sort_free(r);
  size_t bytes = 0;
  VX_HIP(vx_sort_pairs_u32(nullptr, &bytes, nullptr, nullptr, nullptr, nullptr, n, kQuerySortBits, st));
  int rc = dalloc(&r.keys, 2 * (size_t)n);
  if (!rc) rc = dalloc(&r.vals, 2 * (size_t)n);
  if (!rc && hipMalloc(&r.temp, std::max<size_t>(bytes, 1)) != hipSuccess)
    rc = vx_fail(VX_E_NOMEM, "ray sort: %zu bytes of sort storage", bytes);
  if (rc) {
    ray_sort_free(r);
    return rc;
  }
  r.temp_bytes = std::max<size_t>(bytes, 1);
  r.cap = n;
  return VX_OK;
}
// The keys of rays[0, n) (scene camera cam_dev, on the device) and their sort:
// the order is then r.vals[n, 2n).
int ray_sort_run(RaySort& r, const vx_camera* cam_dev, const vx_ray* rays, uint32_t n, hipStream_t st) {
  VX_HIP(vx_launch_query_keys(cam_dev, rays, n, r.keys, r.vals, st));
  size_t bytes = r.temp_bytes;
  VX_HIP(vx_sort_pairs_u32(r.temp, &bytes, r.keys, r.keys + n, r.vals, r.vals + n, n, kQuerySortBits, st));
  return VX_OK;
}

int vx_batch_set_rays(vx_batch* b, const vx_camera* scene_cam, const vx_ray* rays_dev, uint32_t n, vx_hit* hits_dev) {
  if (!b || (n && (!scene_cam || !rays_dev || !hits_dev))) return vx_fail(VX_E_INVALID, "vx_batch_set_rays: null argument");
  DeviceGuard g(b->ctx->device);
  if (!b->rays_dev) {
    int rc = dalloc(&b->rays_dev, 1);
    if (rc) return rc;
    void* pin = nullptr;
    if (hipHostMalloc(&pin, sizeof(vxd::BatchRays), hipHostMallocDefault) != hipSuccess || !pin)
      return vx_fail(VX_E_NOMEM, "vx_batch_set_rays: pinned host allocation");
    b->rays_pin = static_cast<vxd::BatchRays*>(pin);
    std::memset(b->rays_pin, 0, sizeof(vxd::BatchRays));
    if (hipEventCreateWithFlags(&b->rays_ev, hipEventDisableTiming) != hipSuccess)
      return vx_fail(VX_E_DEVICE, "vx_batch_set_rays: hipEventCreate");
  }
  // The last upload must have read the pinned copy before it is rewritten.
  VX_HIP(hipEventSynchronize(b->rays_ev));
  vxd::BatchRays& r = *b->rays_pin;
  std::memset(&r, 0, sizeof(r));
  b->rays_sort = n >= kQuerySortMin && b->ctx->qcfg.sort_rays != 0u;
  if (b->rays_sort && n > b->rsort.cap) {
    VX_HIP(hipEventSynchronize(b->ev[1]));  // the last render, on whatever stream, is done with the buffers
    if (int rc = ray_sort_reserve(b->rsort, n, b->ctx->stream)) return rc;
  }
  if (n) {
    r.cam = *scene_cam;
    r.rays = rays_dev;
    r.hits = hits_dev;
    r.order = b->rays_sort ? b->rsort.vals + n : nullptr;
    r.n = n;
    r.refill_below = b->ctx->qcfg.refill_below;
  }
  b->rays_n = n;
  b->rays_dirty = true;
  return VX_OK;
}

int vx_batch_set_config(vx_batch* b, const vx_batch_config* cfg) {
  if (!b || !cfg) return vx_fail(VX_E_INVALID, "vx_batch_set_config: null argument");
  if (cfg->waves_per_simd < 1u || cfg->waves_per_simd > 16u || cfg->grab < 1u || cfg->grab > 64u ||
      cfg->tile_shape >= vxd::kTileShapes)
    return vx_fail(VX_E_INVALID, "vx_batch_set_config: waves_per_simd in 1..16, grab in 1..64, tile_shape in 0..2");
  b->cfg = *cfg;
  b->B.grab = cfg->grab;
  return VX_OK;
}

int vx_batch_get_config(const vx_batch* b, vx_batch_config* cfg) {
  if (!b || !cfg) return vx_fail(VX_E_INVALID, "vx_batch_get_config: null argument");
  *cfg = b->cfg;
  return VX_OK;
}

int vx_batch_render(vx_batch* b, uint32_t mask, void* stream) {
  if (!b) return vx_fail(VX_E_INVALID, "vx_batch_render: null batch");
  if (mask & (VX_PASS_WRITE_HITS | VX_PASS_COUNT)) return vx_fail(VX_E_INVALID, "vx_batch_render: hits and counting run per view");
  vx_ctx* c = b->ctx;
  DeviceGuard g(c->device);
  hipStream_t st = stream ? (hipStream_t)stream : c->stream;
  const uint32_t n = (uint32_t)b->views.size();
  bool ping_changed = false;
  for (uint32_t i = 0; i < n; ++i) {
    if (b->views[i]->gen != b->gens[i]) return vx_fail(VX_E_INVALID, "vx_batch_render: a view was resized; recreate the batch");
    if (b->ping_host[i] != (uint32_t)b->views[i]->ping) {
      b->ping_host[i] = (uint32_t)b->views[i]->ping;
      ping_changed = true;
    }
  }
  if (ping_changed) {  // a view was rendered on its own since the last batch
    VX_HIP(hipMemcpyAsync(b->ping, b->ping_host.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, st));
    VX_HIP(hipStreamSynchronize(st));
  }
  const vxd::Scene S = make_scene(c, false), SG = make_scene(c, true);
  // Ray queries ride in the primary launch: it then keeps every persistent
  // workgroup (they trace rays once the tiles run out).
  const bool rays = b->rays_n != 0u && (mask & VX_PASS_PRIMARY);
  vxd::Batch B = b->B;
  B.rays = rays ? b->rays_dev : nullptr;
  B.shape = b->cfg.tile_shape;
  if (B.shape != 0u && b->lists_cap < B.tiles) {
    dfree(b->lists);
    b->lists = nullptr;
    b->lists_cap = 0;
    drop_graphs(b);  // captured launches hold the old buffer
    if (int rc = dalloc(&b->lists, (size_t)B.tiles * vxd::kListWords)) return rc;
    b->lists_cap = B.tiles;
  }
  B.lists = b->lists;
  const uint32_t exec_tiles = B.shape == 0u ? B.tiles : B.xtiles[B.shape - 1u];
  const uint32_t persistent = rays ? persistent_groups(b->cfg) : std::min(persistent_groups(b->cfg), std::max(exec_tiles, 1u));
  if (rays && b->rays_dirty) {
    VX_HIP(hipMemcpyAsync(b->rays_dev, b->rays_pin, sizeof(vxd::BatchRays), hipMemcpyHostToDevice, st));
    VX_HIP(hipEventRecord(b->rays_ev, st));
    b->rays_dirty = false;
  }
  if (mask & VX_PASS_PRIMARY) {
    const int rc = refresh_chunk_lights(c, st);
    if (rc) return rc;
  }
  bool profile = false;
  for (const vx_view* v : b->views) profile |= (v->profile_flags & 1u) != 0u;
  VX_HIP(hipEventRecord(b->ev[0], st));
  if (rays && b->rays_sort) {
    // the fused rays' order (keys read the uploaded descriptor's camera), every render: the caller may
    // have rewritten the rays since the last one
    if (int rc = ray_sort_run(b->rsort, &b->rays_dev->cam, b->rays_pin->rays, b->rays_n, st)) return rc;
  }
  if (b->cfg.use_graph) {
    const uint64_t key = ((uint64_t)rays << 63) | ((uint64_t)profile << 62) | ((uint64_t)mask << 32) |
                         ((uint64_t)b->cfg.tile_shape << 24) | ((uint64_t)b->cfg.waves_per_simd << 16) | b->cfg.grab;
    auto it = b->graphs.find(key);
    if (it == b->graphs.end()) {
      hipGraph_t graph = nullptr;
      hipGraphExec_t exec = nullptr;
      hipStream_t cap = nullptr;
      VX_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
      hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
      if (e == hipSuccess) {
        hipError_t le = vx_launch_batch(S, SG, B, mask, persistent, profile, cap);
        e = hipStreamEndCapture(cap, &graph);
        if (e == hipSuccess) e = le;
      }
      if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      if (graph) hipGraphDestroy(graph);
      hipStreamDestroy(cap);
      if (e != hipSuccess) return vx_fail(VX_E_DEVICE, "vx_batch_render: graph capture: %s", hipGetErrorString(e));
      it = b->graphs.emplace(key, exec).first;
    }
    VX_HIP(hipGraphLaunch(it->second, st));
  } else {
    VX_HIP(vx_launch_batch(S, SG, B, mask, persistent, profile, st));
  }
  VX_HIP(hipEventRecord(b->ev[1], st));
  for (uint32_t i = 0; i < n; ++i) {
    vx_view* v = b->views[i];
    const int pong = 1 - v->ping;
    if (mask & VX_PASS_PRIMARY) v->written_primary = pong;
    if (mask & VX_PASS_LOCAL_TAA) v->written_local = pong;
    if (mask & VX_PASS_GODRAY) v->written_godray = pong;
    if (mask & VX_PASS_COMPOSITE_TAA) v->written_out = pong;
    v->last_mask = mask;
    if (mask & VX_PASS_END_FRAME) {
      v->ping = pong;
      b->ping_host[i] = (uint32_t)pong;
    }
  }
  return VX_OK;
}

int vx_batch_timing(vx_batch* b, float* ms) {
  if (!b || !ms) return vx_fail(VX_E_INVALID, "vx_batch_timing: null argument");
  DeviceGuard g(b->ctx->device);
  VX_HIP(hipEventSynchronize(b->ev[1]));
  VX_HIP(hipEventElapsedTime(ms, b->ev[0], b->ev[1]));
  return VX_OK;
}

int vx_render_views_noalloc(vx_ctx* c, vx_view* const* views, uint32_t n, uint32_t mask, void* stream) {
  if (!c || (!views && n)) return vx_fail(VX_E_INVALID, "vx_render_views_noalloc: null argument");
  if (n > VX_NOALLOC_MAX_VIEWS)
    return vx_fail(VX_E_CAPACITY, "vx_render_views_noalloc: %u views, at most %u (the static scratch's size)", n,
                   (unsigned)VX_NOALLOC_MAX_VIEWS);
  if (mask & (VX_PASS_WRITE_HITS | VX_PASS_COUNT))
    return vx_fail(VX_E_INVALID, "vx_render_views_noalloc: hits and counting run per view");
  if (!n) return VX_OK;
  if (int rc = check_batch_views(c, views, n, "vx_render_views_noalloc")) return rc;
  DeviceGuard g(c->device);
  NoAllocScratch* sc = (c->device >= 0 && c->device < kMaxDevices) ? g_scratch_ptr[c->device] : nullptr;
  if (!sc) return vx_fail(VX_E_DEVICE, "vx_render_views_noalloc: the static scratch of device %d is not resolved", c->device);
  hipStream_t st = stream ? (hipStream_t)stream : c->stream;
  // Table order: largest render first (stable), in fixed-size host arrays.
  vx_view* order[VX_NOALLOC_MAX_VIEWS];
  for (uint32_t i = 0; i < n; ++i) order[i] = views[i];
  for (uint32_t i = 1; i < n; ++i)
    for (uint32_t j = i; j > 0 && larger_render(order[j], order[j - 1]); --j) std::swap(order[j], order[j - 1]);
  vxd::BatchView tab[VX_NOALLOC_MAX_VIEWS];
  uint32_t ping[VX_NOALLOC_MAX_VIEWS];
  vxd::Batch B{};
  batch_entries(order, n, tab, B.tiles, B.rblks, B.pblks);
  bool profile = false;
  for (uint32_t i = 0; i < n; ++i) {
    ping[i] = (uint32_t)order[i]->ping;
    profile |= (order[i]->profile_flags & 1u) != 0u;
  }
  // The copies read host memory: wait for them before the host arrays go away.
  VX_HIP(hipMemcpyAsync(sc->table, tab, n * sizeof(vxd::BatchView), hipMemcpyHostToDevice, st));
  VX_HIP(hipMemcpyAsync(sc->ping, ping, n * sizeof(uint32_t), hipMemcpyHostToDevice, st));
  VX_HIP(hipStreamSynchronize(st));
  B.views = sc->table;
  B.ping = sc->ping;
  B.cursor = sc->cursor;   // cleared by the launch's first kernel
  B.n = n;
  B.grab = 1u;
  if (mask & VX_PASS_PRIMARY) {
    const int rc = refresh_chunk_lights(c, st);
    if (rc) return rc;
  }
  const uint32_t persistent = std::min(persistent_groups(vx_batch_config{6u, 1u, 0u, 0u}), std::max(B.tiles, 1u));
  VX_HIP(vx_launch_batch(make_scene(c, false), make_scene(c, true), B, mask, persistent, profile, st));
  for (uint32_t i = 0; i < n; ++i) {
    vx_view* v = order[i];
    const int pong = 1 - v->ping;
    if (mask & VX_PASS_PRIMARY) v->written_primary = pong;
    if (mask & VX_PASS_LOCAL_TAA) v->written_local = pong;
    if (mask & VX_PASS_GODRAY) v->written_godray = pong;
    if (mask & VX_PASS_COMPOSITE_TAA) v->written_out = pong;
    v->last_mask = mask;
    if (mask & VX_PASS_END_FRAME) v->ping = pong;
  }
  return VX_OK;
}

int vx_batch_tune(vx_batch* b, uint32_t mask, uint32_t iters, vx_batch_config* chosen) {
  if (!b) return vx_fail(VX_E_INVALID, "vx_batch_tune: null batch");
  mask &= ~(uint32_t)VX_PASS_END_FRAME;  // re-running the passes without the flip leaves the history as it was
  if (!(mask & (VX_PASS_PRIMARY | VX_PASS_GODRAY))) return vx_fail(VX_E_INVALID, "vx_batch_tune: no work-stealing pass in mask");
  iters = std::max(1u, iters);
  const vx_batch_config keep = b->cfg;
  vx_batch_config best = keep;
  float best_ms = 1e30f;
  const uint32_t waves[] = {4u, 8u};
  const uint32_t grabs[] = {1u, 2u};
  for (uint32_t shape = 0; shape < vxd::kTileShapes; ++shape)
  for (uint32_t w : waves)
    for (uint32_t gsz : grabs) {
      vx_batch_config cfg{w, gsz, keep.use_graph, shape};
      int rc = vx_batch_set_config(b, &cfg);
      if (rc) return rc;
      float total = 0.0f;
      for (uint32_t it = 0; it <= iters; ++it) {
        if ((rc = vx_batch_render(b, mask, nullptr))) return rc;
        float ms = 0.0f;
        if ((rc = vx_batch_timing(b, &ms))) return rc;
        if (it > 0) total += ms;  // the first run warms up (and captures the graph)
      }
      if (total < best_ms) {
        best_ms = total;
        best = cfg;
      }
    }
  vx_batch_set_config(b, &best);
  if (chosen) *chosen = best;
  return VX_OK;
}

int vx_sync(vx_ctx* ctx) {
  if (!ctx) return vx_fail(VX_E_INVALID, "vx_sync: null context");
  DeviceGuard g(ctx->device);
  VX_HIP(hipStreamSynchronize(ctx->stream));
  return VX_OK;
}

int vx_query_rays(vx_ctx* ctx, const vx_camera* scene_cam, const vx_ray* rays_dev, uint32_t n, vx_hit* hits_dev, void* stream) {
  if (!ctx || !scene_cam || (n && (!rays_dev || !hits_dev))) return vx_fail(VX_E_INVALID, "vx_query_rays: null argument");
  DeviceGuard g(ctx->device);
  hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
  if (!ctx->qcam_dev) {
    int rc = dalloc(&ctx->qcam_dev, 1);
    if (!rc) rc = dalloc(&ctx->qcursor, 1);
    for (int i = 0; i < 2 && !rc; ++i)
      if (hipEventCreate(&ctx->qev[i]) != hipSuccess) rc = vx_fail(VX_E_DEVICE, "vx_query_rays: hipEventCreate");
    if (rc) return rc;
  }
  VX_HIP(hipMemcpyAsync(ctx->qcam_dev, scene_cam, sizeof(vx_camera), hipMemcpyHostToDevice, st));
  VX_HIP(hipEventRecord(ctx->qev[0], st));
  if (n) {
    const vx_batch_config wave_cfg{ctx->qcfg.waves_per_simd, 1u, 0u, 0u};
    // never more persistent workgroups than 64-ray blocks
    const uint32_t groups = std::min(persistent_groups(wave_cfg), (n + 63u) / 64u);
    const uint32_t* order = nullptr;
    if (ctx->qcfg.sort_rays && n >= kQuerySortMin) {
      // Trace in the order of the rays' sort keys (direction, then origin).
      if (int rc = ray_sort_reserve(ctx->qsort, n, st)) return rc;
      if (int rc = ray_sort_run(ctx->qsort, ctx->qcam_dev, rays_dev, n, st)) return rc;
      order = ctx->qsort.vals + n;
    }
    VX_HIP(vx_launch_query_ordered(make_scene(ctx, false), ctx->qcam_dev, rays_dev, hits_dev, n, order, ctx->qcursor,
                                   groups, ctx->qcfg.refill_below, st));
  }
  VX_HIP(hipEventRecord(ctx->qev[1], st));
  return VX_OK;
}

int vx_query_count(vx_ctx* ctx, const vx_camera* scene_cam, const vx_ray* rays_dev, uint32_t n, vx_hit* hits_dev,
                   int executed, uint64_t out[9], uint64_t parts[VX_COUNT_PARTS]) {
  if (!ctx || !scene_cam || !out || (n && (!rays_dev || !hits_dev))) return vx_fail(VX_E_INVALID, "vx_query_count: null argument");
  DeviceGuard g(ctx->device);
  if (!ctx->qcam_dev) {
    int rc = dalloc(&ctx->qcam_dev, 1);
    if (!rc) rc = dalloc(&ctx->qcursor, 1);
    for (int i = 0; i < 2 && !rc; ++i)
      if (hipEventCreate(&ctx->qev[i]) != hipSuccess) rc = vx_fail(VX_E_DEVICE, "vx_query_count: hipEventCreate");
    if (rc) return rc;
  }
  if (!ctx->qcounters) {
    const int rc = dalloc(&ctx->qcounters, kViewCounters);
    if (rc) return rc;
  }
  VX_HIP(hipStreamSynchronize(ctx->stream));
  VX_HIP(hipMemset(ctx->qcounters, 0, kViewCounters * sizeof(unsigned long long)));
  VX_HIP(hipMemcpy(ctx->qcam_dev, scene_cam, sizeof(vx_camera), hipMemcpyHostToDevice));
  VX_HIP(vx_launch_query_count(make_scene(ctx, false), ctx->qcam_dev, rays_dev, hits_dev, n, executed != 0, ctx->qcounters,
                               ctx->stream));
  VX_HIP(hipStreamSynchronize(ctx->stream));
  unsigned long long tmp[kViewCounters];
  VX_HIP(hipMemcpy(tmp, ctx->qcounters, sizeof(tmp), hipMemcpyDeviceToHost));
  for (int i = 0; i < 9; ++i) out[i] = tmp[i];
  if (parts)
    for (int i = 0; i < VX_COUNT_PARTS; ++i) parts[i] = tmp[vxd::kNumCounts + i];
  return VX_OK;
}

int vx_query_timing(vx_ctx* ctx, float* ms) {
  if (!ctx || !ms) return vx_fail(VX_E_INVALID, "vx_query_timing: null argument");
  if (!ctx->qev[1]) return vx_fail(VX_E_INVALID, "vx_query_timing: no query has run");
  DeviceGuard g(ctx->device);
  VX_HIP(hipEventSynchronize(ctx->qev[1]));
  VX_HIP(hipEventElapsedTime(ms, ctx->qev[0], ctx->qev[1]));
  return VX_OK;
}

int vx_query_set_config(vx_ctx* ctx, const vx_query_config* cfg) {
  if (!ctx || !cfg) return vx_fail(VX_E_INVALID, "vx_query_set_config: null argument");
  if (cfg->waves_per_simd < 1u || cfg->waves_per_simd > 16u || cfg->refill_below < 1u || cfg->refill_below > 64u ||
      cfg->sort_rays > 1u)
    return vx_fail(VX_E_INVALID, "vx_query_set_config: waves_per_simd in 1..16, refill_below in 1..64, sort_rays 0 or 1");
  ctx->qcfg = *cfg;
  return VX_OK;
}

int vx_query_get_config(const vx_ctx* ctx, vx_query_config* cfg) {
  if (!ctx || !cfg) return vx_fail(VX_E_INVALID, "vx_query_get_config: null argument");
  *cfg = ctx->qcfg;
  return VX_OK;
}

int vx_query_tune(vx_ctx* ctx, const vx_camera* scene_cam, const vx_ray* rays_dev, uint32_t n, vx_hit* hits_dev,
                  vx_query_config* chosen) {
  if (!ctx) return vx_fail(VX_E_INVALID, "vx_query_tune: null context");
  const uint32_t waves[] = {4u, 6u, 8u};
  const uint32_t refill[] = {16u, 32u, 48u, 64u};
  vx_query_config best = ctx->qcfg;
  float best_ms = 1e30f;
  for (uint32_t w : waves)
    for (uint32_t r : refill) {
      vx_query_config cfg{w, r, ctx->qcfg.sort_rays, 0u};  // the sort switch stays as it was set
      int rc = vx_query_set_config(ctx, &cfg);
      float ms = 0.0f;
      if (!rc) rc = vx_query_rays(ctx, scene_cam, rays_dev, n, hits_dev, nullptr);
      if (!rc) rc = vx_query_timing(ctx, &ms);
      if (rc) return rc;
      if (ms < best_ms) {
        best_ms = ms;
        best = cfg;
      }
    }
  ctx->qcfg = best;
  if (chosen) *chosen = best;
  return VX_OK;
}

}  // extern "C"

int vx_internal_batch_rebind(vx_batch* b, vx_view* const* views, uint32_t n) {
  if (!b || !views || !n) return vx_fail(VX_E_INVALID, "vx_batch rebind: bad argument");
  int rc = check_batch_views(b->ctx, views, n, "vx_batch rebind");
  if (rc) return rc;
  DeviceGuard g(b->ctx->device);
  return batch_bind(b, views, n);
}
