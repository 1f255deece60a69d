// vx_server.hip — the serving loop of libvoxhip (vx_server_* in voxhip.h).
//
// No reference counterpart: the reference renders one window's view per
// frame.  This is the continuous-batching front end over the same kernels:
//   * requests (a frame of a client's view stream, or a batch of ray queries)
//     are admitted against a device-memory budget at submit time and queued;
//   * one device thread repeatedly takes everything queued, renders all
//     pending view frames as ONE fused batch (vx_batch_render) on the
//     context stream while all pending ray batches run as ONE fused query
//     launch on a second stream, so the two overlap on the GPU; requests that
//     arrive meanwhile wait for the next launch;
//   * each view stream keeps its vx_view (history) resident between frames;
//   * edits rebuild their chunk on host worker threads (fill + edits + the
//     chunk builder); at the next launch boundary each rebuilt chunk goes to a
//     fresh node-arena range (best-fit allocator), then its macro / colinfo /
//     meta entries are rewritten and the old range is freed: the arena holds
//     old and new copies until the swap (double-buffered replacement), and no
//     launch is in flight at the boundary;
//   * with VX_SERVER_DEVICE_BUILD the chunks' material stays resident in HBM
//     instead: at a launch boundary the pending edits are scattered into it
//     and the edited chunks are rebuilt by the device chunk builder straight
//     into fresh arena ranges, then the same meta swap.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "vx_internal.h"

namespace {

double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

constexpr uint64_t kViewBytesPerPx = 250;  // history + outputs + hits of one view (see vx_view_create)
constexpr uint64_t kRayBytes = sizeof(vx_ray) + sizeof(vx_hit);
// With a latency target, one launch takes at most this share of the target's
// worth of rays at the measured rate: a request that arrives just after a
// launch starts waits for it and then runs in the next, so two launches must
// fit in the target.
constexpr double kLaunchShare = 1.0 / 3.0;

// Best-fit allocator over the node arena (element units).
class Arena {
 public:
  void init(uint64_t cap, const std::vector<std::pair<uint64_t, uint64_t>>& used) {
    free_.clear();
    by_size_.clear();
    std::vector<std::pair<uint64_t, uint64_t>> u = used;
    std::sort(u.begin(), u.end());
    uint64_t at = 0;
    for (auto& [off, len] : u) {
      if (off > at) add(at, off - at);
      at = std::max(at, off + len);
    }
    if (cap > at) add(at, cap - at);
  }
  bool alloc(uint64_t len, uint64_t& off) {
    if (len == 0) len = 1;
    auto it = by_size_.lower_bound(len);
    if (it == by_size_.end()) return false;
    off = it->second;
    const uint64_t size = it->first;
    by_size_.erase(it);
    free_.erase(off);
    if (size > len) add(off + len, size - len);
    return true;
  }
  void release(uint64_t off, uint64_t len) {
    if (len == 0) len = 1;
    auto next = free_.lower_bound(off);
    if (next != free_.end() && next->first == off + len) {  // merge with the following block
      len += next->second;
      erase(next->first, next->second);
    }
    auto prev = free_.lower_bound(off);
    if (prev != free_.begin()) {
      --prev;
      if (prev->first + prev->second == off) {  // merge with the preceding block
        off = prev->first;
        len += prev->second;
        erase(prev->first, prev->second);
      }
    }
    add(off, len);
  }
  uint64_t free_total() const {
    uint64_t t = 0;
    for (auto& kv : free_) t += kv.second;
    return t;
  }

 private:
  void add(uint64_t off, uint64_t len) {
    free_[off] = len;
    by_size_.emplace(len, off);
  }
  void erase(uint64_t off, uint64_t len) {
    free_.erase(off);
    auto range = by_size_.equal_range(len);
    for (auto it = range.first; it != range.second; ++it)
      if (it->second == off) {
        by_size_.erase(it);
        break;
      }
  }
  std::map<uint64_t, uint64_t> free_;            // offset -> length
  std::multimap<uint64_t, uint64_t> by_size_;    // length -> offset
};

// Rays of one launch or queue, by kind: view pixels (primary, godray,
// composite and TAA per pixel) and query rays (a primary trace only) cost
// different amounts each.
struct Work {
  double view = 0.0, query = 0.0;
  bool empty() const { return view <= 0.0 && query <= 0.0; }
  Work& operator+=(const Work& o) {
    view += o.view;
    query += o.query;
    return *this;
  }
  Work& operator-=(const Work& o) {
    view -= o.view;
    query -= o.query;
    return *this;
  }
};

// Launch time as fixed + per_view * view pixels + per_query * query rays,
// fitted over recent launches by exponentially weighted least squares (weight
// 0.9 per launch, about the last ten).  A launch's fixed costs (launch
// boundary, camera writes, syncs) are not spread over its rays, and the two
// kinds get their own rate.  While the launches seen so far cannot separate
// the three terms (one kind only, or one mix), the missing terms fall back to
// the fit on all rays (fixed + per_ray * rays), and that to proportional.
class LaunchCost {
 public:
  void add(const Work& w, double ms) {
    const double x[3] = {1.0, w.view, w.query};
    for (int i = 0; i < 3; ++i) {
      for (int j = 0; j < 3; ++j) a_[i][j] = kDecay * a_[i][j] + x[i] * x[j];
      b_[i] = kDecay * b_[i] + x[i] * ms;
    }
    fit();
  }
  bool ready() const { return a_[0][0] > 0.5 && (a_[1][1] > 0.0 || a_[2][2] > 0.0); }
  double fixed() const { return c_[0]; }
  double ms(const Work& w) const { return c_[0] + c_[1] * w.view + c_[2] * w.query; }
  // Time of `w` beyond the fixed cost.
  double rays_ms(const Work& w) const { return c_[1] * w.view + c_[2] * w.query; }

 private:
  void fit() {
    const double n = a_[0][0];
    // All rays as one kind: fixed + per_ray * (view + query).
    const double sx = a_[0][1] + a_[0][2], sy = b_[0];
    const double sxx = a_[1][1] + 2.0 * a_[1][2] + a_[2][2], sxy = b_[1] + b_[2];
    const double mx = sx / n, my = sy / n, var = sxx / n - mx * mx;
    double per = var > 1e-6 * mx * mx ? (sxy / n - mx * my) / var : 0.0;
    if (!(per > 0.0)) per = my / std::max(mx, 1.0);
    double fixed = std::min(std::max(my - per * mx, 0.0), my);
    c_[0] = fixed;
    c_[1] = c_[2] = per;
    // Three terms, when both kinds vary independently enough to separate
    // them: the normal equations in variables scaled to unit mean square, so
    // a small pivot means the launches seen are (nearly) collinear.
    if (!(a_[1][1] > 0.0 && a_[2][2] > 0.0)) return;
    const double sc[3] = {1.0, std::sqrt(n / a_[1][1]), std::sqrt(n / a_[2][2])};
    double m[3][4];
    for (int i = 0; i < 3; ++i) {
      for (int j = 0; j < 3; ++j) m[i][j] = a_[i][j] / n * sc[i] * sc[j];
      m[i][3] = b_[i] / n * sc[i];
    }
    for (int k = 0; k < 3; ++k) {  // Gauss-Jordan elimination with partial pivoting
      int piv = k;
      for (int i = k + 1; i < 3; ++i)
        if (std::fabs(m[i][k]) > std::fabs(m[piv][k])) piv = i;
      if (std::fabs(m[piv][k]) < 1e-4) return;  // not separable yet: keep the fit on all rays
      for (int j = 0; j < 4; ++j) std::swap(m[k][j], m[piv][j]);
      for (int i = 0; i < 3; ++i) {
        if (i == k) continue;
        const double f = m[i][k] / m[k][k];
        for (int j = k; j < 4; ++j) m[i][j] -= f * m[k][j];
      }
    }
    const double f3 = m[0][3] / m[0][0], v3 = m[1][3] / m[1][1] * sc[1], q3 = m[2][3] / m[2][2] * sc[2];
    if (!(f3 >= 0.0 && v3 > 0.0 && q3 > 0.0)) return;
    c_[0] = f3;
    c_[1] = v3;
    c_[2] = q3;
  }
  static constexpr double kDecay = 0.9;
  double a_[3][3] = {}, b_[3] = {};
  double c_[3] = {};  // fixed, per view pixel, per query ray
};

// Admission multiplies its estimate by a margin before comparing it with the
// target: the 99th percentile of (latency / estimate) over the last kCalib
// requests admitted near the threshold (estimate at least half the admissible
// one), so that p99 latency holds the target whatever the model misses
// (launches that later arrivals fill up, scenes that cost more per ray);
// kAdmitMargin until kCalibMin such requests have finished.
constexpr double kAdmitMargin = 1.1;
constexpr size_t kCalib = 256, kCalibMin = 32;
constexpr double kMarginMax = 4.0;
// An idle server admits the next request unconditionally after this long without a launch.
constexpr double kProbeAfterMs = 100.0;

constexpr uint64_t kNoPin = ~0ull;

struct Pending {
  vx_request req;            // req.rays is the caller's and not read after submit
  uint64_t pin = kNoPin;     // RAYS: element offset of its block in the pinned ray/hit pool
  std::vector<vx_ray> rays;  // RAYS without a pool block: copied at submit
  double submit_ms = 0.0;
  double est_ms = -1.0;      // admission's latency estimate (without margin), -1 if admitted without one
  uint64_t reserved = 0;     // budget bytes released at completion
};

struct Stream {
  vx_view* view = nullptr;
  uint32_t w = 0, h = 0;
  uint64_t bytes = 0;
  bool busy = false;  // a frame of this stream is queued or in flight
};

struct ChunkBuild {
  uint32_t slot = 0;
  uint64_t gen = 0;  // the slot's edit generation this build includes
  std::vector<vx_node> nodes;
  std::vector<vx_node_ropes> ropes;
  uint32_t macro[16];
  uint32_t colinfo[2048];
};

}  // namespace

struct vx_server {
  vx_ctx* ctx = nullptr;
  vx_server_config cfg{};
  hipStream_t qstream = nullptr;
  // world
  std::vector<vx_chunk_meta> meta;  // host mirror of the device meta table (device thread only after create)
  std::vector<std::array<int32_t, 3>> origin;  // per slot: chunk origin in voxels (immutable, read by workers)
  Arena arena;
  std::unordered_map<uint32_t, std::unordered_map<uint32_t, uint8_t>> edits;  // slot -> voxel index -> material
  std::vector<uint8_t> dirty;        // slot queued for rebuild
  std::vector<uint64_t> gen;         // per slot: edit generation (bumped by each queued edit batch)
  std::vector<uint64_t> applied;     // per slot: generation of the build on the device
  uint32_t building = 0;             // builds taken by workers and not yet in `ready`
  std::deque<uint32_t> rebuild_q;
  std::vector<ChunkBuild> ready;     // rebuilt, waiting for a launch boundary
  // Rebuilt chunks the node arena could not place at the last boundary.  They
  // are retried at the next boundary that something else starts; the device
  // thread does not wake for them alone (the arena changes only when other
  // builds are applied, so retrying on their own would spin).
  std::vector<ChunkBuild> deferred;
  std::vector<std::thread> workers;
  std::condition_variable wcv;
  // requests
  std::mutex mu;
  std::condition_variable cv;
  std::deque<Pending> inbox;
  std::deque<vx_result> outbox;
  std::unordered_map<uint32_t, Stream> streams;
  uint64_t mem_in_use = 0;
  // Latency target: the launch-time model, rays admitted and not yet
  // launched, and the launch in flight (its rays and start time).
  LaunchCost cost;
  Work queued, inflight;
  double margin = kAdmitMargin;
  std::vector<double> calib;  // ring of latency / estimate ratios
  size_t calib_at = 0;
  double last_launch_end_ms = 0.0;
  double inflight_t0 = 0.0;
  bool stop = false;
  std::thread device;
  // stats
  vx_server_stats st{};
  std::vector<double> lat_view, lat_rays;
  double t_start = 0.0;
  int last_error = VX_OK;
  // Device builds (VX_SERVER_DEVICE_BUILD): per-slot material (64^3 B) and
  // ground (64^2 i32) resident in HBM, the builder, pending edits (slot ->
  // voxel index -> material, the last edit of a voxel wins), dirty slots in
  // edit order, and device buffers for the edit scatter.
  bool device_build = false;
  uint8_t* dmat = nullptr;
  int32_t* dground = nullptr;
  vx_chunk_builder* builder = nullptr;
  uint32_t build_batch = 0;
  std::unordered_map<uint32_t, std::unordered_map<uint32_t, uint8_t>> pend;
  std::deque<uint32_t> dev_q;
  std::vector<uint32_t> dev_deferred;  // device builds the arena could not place (as `deferred`)
  bool dev_busy = false;
  uint64_t* d_off = nullptr;
  uint8_t* d_val = nullptr;
  size_t d_edit_cap = 0;
  // Pinned pool for ray batches: a request's rays are copied into its block
  // on the submitting thread and its hits come back into the block at the
  // same offset, so a launch queues only DMA copies.  Requests that find no
  // room go through the launch's own staging.
  vx_ray* pin_rays = nullptr;
  vx_hit* pin_hits = nullptr;
  Arena pin;  // element offsets (guarded by mu)
};

namespace {

uint32_t slot_of_voxel(const vx_server* s, int32_t wx, int32_t wy, int32_t wz, uint32_t& idx) {
  const int32_t cx = (int32_t)std::floor(wx / 64.0), cy = (int32_t)std::floor(wy / 64.0), cz = (int32_t)std::floor(wz / 64.0);
  const int32_t ix = cx - s->cfg.grid_origin[0], iy = cy - s->cfg.grid_origin[1], iz = cz - s->cfg.grid_origin[2];
  if (ix < 0 || iy < 0 || iz < 0 || ix >= (int32_t)s->cfg.grid_dims[0] || iy >= (int32_t)s->cfg.grid_dims[1] ||
      iz >= (int32_t)s->cfg.grid_dims[2])
    return VX_INVALID_U32;
  // Slots follow grid order (x fastest), as vx_host_world packs them.
  const uint32_t slot = ((uint32_t)iz * s->cfg.grid_dims[1] + (uint32_t)iy) * s->cfg.grid_dims[0] + (uint32_t)ix;
  if (slot >= s->meta.size()) return VX_INVALID_U32;
  const int32_t lx = wx - cx * 64, ly = wy - cy * 64, lz = wz - cz * 64;
  idx = (uint32_t)(ly * 4096 + lz * 64 + lx);  // material layout idx = y*4096 + z*64 + x
  return slot;
}

// Host worker: rebuild queued chunks (fill, apply the chunk's edits, build).
void worker_main(vx_server* s) {
  std::vector<uint8_t> mat(64 * 64 * 64);
  std::vector<int32_t> ground(64 * 64);
  ChunkBuild b;
  b.nodes.resize(VX_MAX_CHUNK_NODES);
  b.ropes.resize(VX_MAX_CHUNK_NODES);
  for (;;) {
    uint32_t slot;
    std::vector<std::pair<uint32_t, uint8_t>> ed;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->wcv.wait(lk, [&] { return s->stop || !s->rebuild_q.empty(); });
      if (s->stop) return;
      slot = s->rebuild_q.front();
      s->rebuild_q.pop_front();
      s->dirty[slot] = 0;  // edits arriving from now on queue the chunk again
      b.gen = s->gen[slot];
      ++s->building;
      for (auto& kv : s->edits[slot]) ed.emplace_back(kv.first, kv.second);
    }
    const std::array<int32_t, 3>& origin = s->origin[slot];
    const int32_t cx = origin[0] / 64, cy = origin[1] / 64, cz = origin[2] / 64;
    uint32_t count = 0;
    int rc = vx_world_fill_chunk(&s->cfg.world, cx, cy, cz, mat.data(), ground.data());
    if (!rc) {
      for (auto& e : ed) mat[e.first] = e.second;
      rc = vx_build_chunk(mat.data(), ground.data(), cy * 64, s->cfg.world.dirt_depth, b.nodes.data(), b.ropes.data(),
                          VX_MAX_CHUNK_NODES, &count, b.macro, b.colinfo);
    }
    if (rc) {
      {
        std::lock_guard<std::mutex> lk(s->mu);
        --s->building;
        s->last_error = rc;
      }
      s->cv.notify_all();  // vx_server_wait_edits and vx_server_drain return on the error
      continue;
    }
    ChunkBuild out;
    out.slot = slot;
    out.gen = b.gen;
    out.nodes.assign(b.nodes.begin(), b.nodes.begin() + count);
    out.ropes.assign(b.ropes.begin(), b.ropes.begin() + count);
    std::memcpy(out.macro, b.macro, sizeof(out.macro));
    std::memcpy(out.colinfo, b.colinfo, sizeof(out.colinfo));
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->ready.push_back(std::move(out));
      --s->building;
    }
    s->cv.notify_all();
  }
}

// Launch boundary: move rebuilt chunks into fresh arena ranges, then swap
// their meta entries and release the old ranges.  No launch is in flight.
int apply_ready(vx_server* s, std::vector<ChunkBuild>& builds) {
  // Newest build of each slot only: builds finish in any order, and an older
  // one must not replace a newer one.
  std::sort(builds.begin(), builds.end(), [](const ChunkBuild& a, const ChunkBuild& b) {
    return a.slot != b.slot ? a.slot < b.slot : a.gen > b.gen;
  });
  std::vector<ChunkBuild> latest;
  for (ChunkBuild& b : builds)
    if ((latest.empty() || latest.back().slot != b.slot) && b.gen > s->applied[b.slot]) latest.push_back(std::move(b));
  builds.swap(latest);
  std::vector<vx_region> rn, rr, rm, rc, mr;
  std::vector<std::pair<uint64_t, uint64_t>> old;
  std::vector<std::pair<uint32_t, vx_chunk_meta>> swaps;
  std::vector<std::pair<uint32_t, uint64_t>> placed;  // (slot, generation) of the builds that get a range
  std::vector<ChunkBuild> later;
  for (ChunkBuild& b : builds) {
    uint64_t off = 0;
    if (!s->arena.alloc(b.nodes.size(), off)) {  // arena full: retry at the next boundary
      later.push_back(std::move(b));
      continue;
    }
    vx_chunk_meta m = s->meta[b.slot];
    old.emplace_back(m.node_base, m.node_count);
    m.node_base = (uint32_t)off;
    m.node_count = (uint32_t)b.nodes.size();
    bool any = false;
    for (uint32_t w : b.macro) any |= w != 0u;
    m.macro_empty = any ? 0u : 1u;
    swaps.emplace_back(b.slot, m);
    placed.emplace_back(b.slot, b.gen);
    rn.push_back(vx_region{off, b.nodes.size(), b.nodes.data()});
    rr.push_back(vx_region{off, b.ropes.size(), b.ropes.data()});
    rm.push_back(vx_region{m.macro_base, 16, b.macro});
    rc.push_back(vx_region{m.colinfo_base, 2048, b.colinfo});
  }
  int rc0 = VX_OK;
  if (!swaps.empty()) {
    // The new copies first, into ranges no meta entry points at yet; then the
    // meta swap; then the old ranges are free.
    if ((rc0 = vx_upload_regions(s->ctx, VX_ARENA_NODES, rn.data(), (uint32_t)rn.size())) ||
        (rc0 = vx_upload_regions(s->ctx, VX_ARENA_ROPES, rr.data(), (uint32_t)rr.size())) ||
        (rc0 = vx_upload_regions(s->ctx, VX_ARENA_MACRO, rm.data(), (uint32_t)rm.size())) ||
        (rc0 = vx_upload_regions(s->ctx, VX_ARENA_COLINFO, rc.data(), (uint32_t)rc.size())))
      return rc0;
    for (auto& sw : swaps) {
      s->meta[sw.first] = sw.second;
      mr.push_back(vx_region{sw.first, 1, &s->meta[sw.first]});
    }
    if ((rc0 = vx_upload_regions(s->ctx, VX_ARENA_META, mr.data(), (uint32_t)mr.size()))) return rc0;
    for (auto& o : old) s->arena.release(o.first, o.second);
    std::lock_guard<std::mutex> lk(s->mu);
    for (auto& pg : placed) s->applied[pg.first] = std::max(s->applied[pg.first], pg.second);
    s->st.edits_applied += swaps.size();
  }
  if (!later.empty()) {
    // Old ranges released above may make room now: then retry at once (the
    // device thread wakes for `ready`); otherwise wait for a boundary that
    // applies other builds.
    const uint64_t free_total = s->arena.free_total();
    const bool arena_changed = !swaps.empty();
    std::lock_guard<std::mutex> lk(s->mu);
    for (auto& b : later) {
      if (b.nodes.size() > free_total && s->last_error == VX_OK)
        s->last_error = vx_fail(VX_E_CAPACITY, "server: chunk slot %u needs %zu nodes, %llu free in the node arena",
                                b.slot, b.nodes.size(), (unsigned long long)free_total);
      (arena_changed ? s->ready : s->deferred).push_back(std::move(b));
    }
  }
  return VX_OK;
}

// Launch boundary, device builds: scatter the batch's edits into the resident
// material, count the edited chunks, place each in a fresh arena range, emit
// them there, then swap their meta entries and release the old ranges.  When
// the arena cannot place every chunk, the batch waits for the next boundary
// (its edits are already in the material).
int apply_device_builds(vx_server* s, const std::vector<uint32_t>& slots,
                        std::unordered_map<uint32_t, std::unordered_map<uint32_t, uint8_t>>& edits, bool& requeue) {
  requeue = false;
  constexpr uint64_t kVox = 64ull * 64ull * 64ull;
  std::vector<uint64_t> off;
  std::vector<uint8_t> val;
  for (auto& e : edits)
    for (auto& v : e.second) {
      off.push_back((uint64_t)e.first * kVox + v.first);
      val.push_back(v.second);
    }
  vx_internal_arenas ar;
  vx_internal_get_arenas(s->ctx, &ar);
  hipStream_t st = (hipStream_t)ar.stream;
  if (!off.empty()) {
    if (off.size() > s->d_edit_cap) {
      if (s->d_off) hipFree(s->d_off);
      if (s->d_val) hipFree(s->d_val);
      s->d_off = nullptr;
      s->d_val = nullptr;
      s->d_edit_cap = 0;
      const size_t cap = std::max<size_t>(off.size(), 4096);
      if (hipMalloc((void**)&s->d_off, cap * sizeof(uint64_t)) != hipSuccess ||
          hipMalloc((void**)&s->d_val, cap) != hipSuccess)
        return vx_fail(VX_E_NOMEM, "server: edit buffers");
      s->d_edit_cap = cap;
    }
    if (hipMemcpyAsync(s->d_off, off.data(), off.size() * sizeof(uint64_t), hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(s->d_val, val.data(), val.size(), hipMemcpyHostToDevice, st) != hipSuccess)
      return vx_fail(VX_E_DEVICE, "server: edit upload");
    int rc = vx_internal_scatter_u8(s->ctx, s->dmat, s->d_off, s->d_val, (uint32_t)off.size());
    if (rc) return rc;
  }
  const uint32_t n = (uint32_t)slots.size();
  std::vector<int32_t> oy(n);
  for (uint32_t i = 0; i < n; ++i) oy[i] = s->meta[slots[i]].origin[1];
  std::vector<uint32_t> counts(n);
  std::vector<uint8_t> solid(n);
  int rc = vx_internal_builder_count(s->builder, n, s->dmat, s->dground, slots.data(), oy.data(),
                                     s->cfg.world.dirt_depth, counts.data(), solid.data());
  if (rc) return rc;
  std::vector<uint64_t> base(n);
  std::vector<uint32_t> mb(n), cb(n);
  for (uint32_t i = 0; i < n; ++i) {
    if (!s->arena.alloc(counts[i], base[i])) {
      for (uint32_t j = 0; j < i; ++j) s->arena.release(base[j], counts[j]);
      const uint64_t free_total = s->arena.free_total();
      if (counts[i] > free_total) {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->last_error == VX_OK)
          s->last_error = vx_fail(VX_E_CAPACITY, "server: chunk slot %u needs %u nodes, %llu free in the node arena",
                                  slots[i], counts[i], (unsigned long long)free_total);
      }
      requeue = true;
      return VX_OK;
    }
    mb[i] = s->meta[slots[i]].macro_base;
    cb[i] = s->meta[slots[i]].colinfo_base;
  }
  if ((rc = vx_chunk_builder_emit(s->builder, base.data(), mb.data(), cb.data()))) return rc;
  std::vector<std::pair<uint64_t, uint64_t>> old;
  std::vector<vx_region> mr;
  for (uint32_t i = 0; i < n; ++i) {
    vx_chunk_meta& m = s->meta[slots[i]];
    old.emplace_back(m.node_base, m.node_count);
    m.node_base = (uint32_t)base[i];
    m.node_count = counts[i];
    m.macro_empty = solid[i] ? 0u : 1u;
    mr.push_back(vx_region{slots[i], 1, &m});
  }
  if ((rc = vx_upload_regions(s->ctx, VX_ARENA_META, mr.data(), (uint32_t)mr.size()))) return rc;
  for (auto& o : old) s->arena.release(o.first, o.second);
  return VX_OK;
}

// Fills the resident material and ground of every slot from the world
// generator, on `threads` host threads, in batches uploaded to the device.
int fill_device_world(vx_server* s, uint32_t threads) {
  constexpr size_t kVox = 64ull * 64ull * 64ull;
  const uint32_t slots = (uint32_t)s->meta.size(), batch = 128;
  std::vector<uint8_t> mat(batch * kVox);
  std::vector<int32_t> ground((size_t)batch * 4096);
  for (uint32_t s0 = 0; s0 < slots; s0 += batch) {
    const uint32_t nb = std::min(batch, slots - s0);
    std::atomic<uint32_t> next{0};
    std::atomic<int> err{VX_OK};
    auto work = [&]() {
      for (uint32_t i; (i = next.fetch_add(1)) < nb;) {
        const vx_chunk_meta& m = s->meta[s0 + i];
        const int rc = vx_world_fill_chunk(&s->cfg.world, m.origin[0] / 64, m.origin[1] / 64, m.origin[2] / 64,
                                           mat.data() + i * kVox, ground.data() + (size_t)i * 4096);
        if (rc) err.store(rc);
      }
    };
    std::vector<std::thread> pool;
    for (uint32_t t = 1; t < threads; ++t) pool.emplace_back(work);
    work();
    for (auto& th : pool) th.join();
    if (err.load()) return err.load();
    if (hipMemcpy(s->dmat + (size_t)s0 * kVox, mat.data(), nb * kVox, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(s->dground + (size_t)s0 * 4096, ground.data(), (size_t)nb * 4096 * 4, hipMemcpyHostToDevice) != hipSuccess)
      return vx_fail(VX_E_DEVICE, "vx_server_create: material upload");
  }
  return VX_OK;
}

double pct(std::vector<double> v, double p) {
  if (v.empty()) return 0.0;
  std::sort(v.begin(), v.end());
  const size_t i = std::min(v.size() - 1, (size_t)(p / 100.0 * (double)(v.size() - 1) + 0.5));
  return v[i];
}

// The camera fields a ray query reads: voxel size, chunk size and count, grid origin and dims.
bool same_scene(const vx_camera& a, const vx_camera& b) {
  return a.voxel_params[0] == b.voxel_params[0] && a.chunk_size == b.chunk_size && a.chunk_count == b.chunk_count &&
         std::memcmp(a.grid_origin_chunk, b.grid_origin_chunk, 3 * sizeof(int32_t)) == 0 &&
         std::memcmp(a.grid_dims, b.grid_dims, 3 * sizeof(uint32_t)) == 0;
}

uint64_t request_rays(const Pending& p) {
  return p.req.kind == VX_REQ_VIEW ? (uint64_t)p.req.width * p.req.height : (uint64_t)p.req.n_rays;
}

// Publishes a finished request (under mu): its result, latency, budget and pool block.
void complete(vx_server* s, Pending& p, int status, double t0, double t1, std::vector<double>& lat) {
  vx_result r{};
  r.tag = p.req.tag;
  r.kind = p.req.kind;
  r.status = status;
  r.submit_ms = p.submit_ms - s->t_start;
  r.start_ms = t0 - s->t_start;
  r.done_ms = t1 - s->t_start;
  s->outbox.push_back(r);
  lat.push_back(t1 - p.submit_ms);
  if (p.est_ms > 0.0 && p.est_ms * s->margin >= 0.5 * s->cfg.latency_target_ms) {
    const double ratio = (t1 - p.submit_ms) / p.est_ms;
    if (s->calib.size() < kCalib)
      s->calib.push_back(ratio);
    else
      s->calib[s->calib_at++ % kCalib] = ratio;
  }
  s->mem_in_use -= p.reserved;
  s->st.rays += request_rays(p);
  if (p.pin != kNoPin) {
    s->pin.release(p.pin, p.req.n_rays);
    p.pin = kNoPin;
  }
  ++s->st.completed;
}

Work request_work(const Pending& p) {
  Work w;
  (p.req.kind == VX_REQ_VIEW ? w.view : w.query) = (double)request_rays(p);
  return w;
}

void device_main(vx_server* s) {
  hipSetDevice(vx_internal_device(s->ctx));
  // Device ray and hit buffers for a launch, and pinned staging for the
  // requests that found no room in the pool (allocated when first needed):
  // pinned host memory lets the copies run on the DMA engines instead of as
  // staging blits on the compute units.  Device buffers hold 4 M rays from
  // the start, doubled when a launch needs more: reallocating (a
  // synchronising free) inside a launch would hold every request in it.
  vx_ray* d_rays = nullptr;
  vx_hit* d_hits = nullptr;
  size_t d_cap = 0;
  vx_ray* h_rays = nullptr;
  vx_hit* h_hits = nullptr;
  size_t h_cap = 0;
  auto grow_dev = [&](size_t nr) -> int {
    if (nr <= d_cap) return VX_OK;
    const size_t cap = std::max({nr, 2 * d_cap, (size_t)1 << 22});
    if (d_rays) hipFree(d_rays);
    if (d_hits) hipFree(d_hits);
    d_rays = nullptr;
    d_hits = nullptr;
    d_cap = 0;
    if (hipMalloc((void**)&d_rays, cap * sizeof(vx_ray)) != hipSuccess ||
        hipMalloc((void**)&d_hits, cap * sizeof(vx_hit)) != hipSuccess)
      return vx_fail(VX_E_NOMEM, "server: device ray buffers for %zu rays", cap);
    d_cap = cap;
    return VX_OK;
  };
  auto grow_host = [&](size_t nr) -> int {
    if (nr <= h_cap) return VX_OK;
    const size_t cap = std::max(nr, 2 * h_cap);
    if (h_rays) hipHostFree(h_rays);
    if (h_hits) hipHostFree(h_hits);
    h_rays = nullptr;
    h_hits = nullptr;
    h_cap = 0;
    if (hipHostMalloc((void**)&h_rays, cap * sizeof(vx_ray), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&h_hits, cap * sizeof(vx_hit), hipHostMallocDefault) != hipSuccess)
      return vx_fail(VX_E_NOMEM, "server: ray staging for %zu rays", cap);
    h_cap = cap;
    return VX_OK;
  };
  if (grow_dev(1) != VX_OK) {
    std::lock_guard<std::mutex> lk(s->mu);
    s->last_error = VX_E_NOMEM;
  }
  // One batch for the thread's lifetime, re-bound to each launch's views: no
  // allocation, event creation or synchronising free per launch.
  vx_batch* batch = nullptr;
  for (;;) {
    std::vector<Pending> views, rays;
    std::vector<ChunkBuild> builds;
    std::vector<uint32_t> dev_slots;
    std::vector<uint64_t> dev_gens;
    std::unordered_map<uint32_t, std::unordered_map<uint32_t, uint8_t>> dev_edits;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv.wait(lk, [&] { return s->stop || !s->inbox.empty() || !s->ready.empty() || !s->dev_q.empty(); });
      if (s->stop) break;
      builds.swap(s->ready);
      for (ChunkBuild& b : s->deferred) builds.push_back(std::move(b));
      s->deferred.clear();
      s->dev_q.insert(s->dev_q.begin(), s->dev_deferred.begin(), s->dev_deferred.end());  // oldest first
      s->dev_deferred.clear();
      // Device builds: the oldest dirty slots, up to one builder batch, with their pending edits.
      while (!s->dev_q.empty() && dev_slots.size() < s->build_batch) {
        const uint32_t slot = s->dev_q.front();
        s->dev_q.pop_front();
        s->dirty[slot] = 0;
        dev_slots.push_back(slot);
        dev_gens.push_back(s->gen[slot]);
        auto it = s->pend.find(slot);
        if (it != s->pend.end()) {
          dev_edits[slot] = std::move(it->second);
          s->pend.erase(it);
        }
      }
      s->dev_busy = !dev_slots.empty();
      // Queued requests join this launch in arrival order: at most one frame
      // per view stream, max_views_per_launch views and, with a latency
      // target, a third of the target's worth of rays (at least one request); the
      // rest waits for the next launch.
      const double cap_ms = s->cfg.latency_target_ms > 0.0f && s->cost.ready()
                                ? kLaunchShare * s->cfg.latency_target_ms
                                : 0.0;
      Work launch;
      uint64_t query_rays = 0;
      std::deque<Pending> keep;
      std::vector<uint32_t> taken;
      while (!s->inbox.empty()) {
        Pending p = std::move(s->inbox.front());
        s->inbox.pop_front();
        const uint64_t n = request_rays(p);
        Work with = launch;
        with += request_work(p);
        const bool full = cap_ms > 0.0 && !launch.empty() && s->cost.ms(with) > cap_ms;
        if (p.req.kind == VX_REQ_VIEW) {
          const bool dup = std::find(taken.begin(), taken.end(), p.req.stream_id) != taken.end();
          if (full || dup || views.size() >= s->cfg.max_views_per_launch) {
            keep.push_back(std::move(p));
          } else {
            taken.push_back(p.req.stream_id);
            launch = with;
            views.push_back(std::move(p));
          }
        } else if (full || (query_rays > 0 && query_rays + n > d_cap)) {  // ray batches beyond the device buffers wait
          keep.push_back(std::move(p));
        } else {
          launch = with;
          query_rays += n;
          rays.push_back(std::move(p));
        }
      }
      s->inbox.swap(keep);
      s->queued -= launch;
      s->inflight = launch;
      s->inflight_t0 = now_ms();
    }
    // The launch's time, for the requests' start and for the cost model,
    // includes applying the edits at this boundary.
    const double t0 = now_ms();
    int rc = builds.empty() ? VX_OK : apply_ready(s, builds);
    if (!rc && !dev_slots.empty()) {
      bool requeue = false;
      rc = apply_device_builds(s, dev_slots, dev_edits, requeue);
      std::lock_guard<std::mutex> lk(s->mu);
      if (!rc && requeue) {  // the arena is full: these chunks wait for the next boundary
        for (uint32_t slot : dev_slots)
          if (!s->dirty[slot]) {
            s->dirty[slot] = 1;
            s->dev_deferred.push_back(slot);
          }
      } else if (!rc) {
        for (size_t i = 0; i < dev_slots.size(); ++i)
          s->applied[dev_slots[i]] = std::max(s->applied[dev_slots[i]], dev_gens[i]);
        s->st.edits_applied += dev_slots.size();
      }
      s->dev_busy = false;
    }
    // Ray batches first: one fused query launch on the second stream.  Their
    // upload is a copy kernel; issued after the views' persistent primary it
    // would wait for that kernel's workgroups to drain.
    size_t nr = 0, n_staged = 0;
    for (Pending& p : rays) {
      nr += p.req.n_rays;
      if (p.pin == kNoPin) n_staged = nr;  // staging keeps the launch's offsets
    }
    if (!rc && nr) {
      rc = grow_dev(nr);
      if (!rc && n_staged) rc = grow_host(n_staged);
      size_t at = 0;
      for (Pending& p : rays) {
        const size_t n = p.req.n_rays;
        const vx_ray* src = p.pin != kNoPin ? s->pin_rays + p.pin : h_rays + at;
        if (!rc && n && p.pin == kNoPin) std::memcpy(h_rays + at, p.rays.data(), n * sizeof(vx_ray));
        if (!rc && n &&
            hipMemcpyAsync(d_rays + at, src, n * sizeof(vx_ray), hipMemcpyHostToDevice, s->qstream) != hipSuccess)
          rc = vx_fail(VX_E_DEVICE, "server: ray upload");
        at += n;
      }
      // One fused launch per run of consecutive requests that share the scene
      // fields of their camera (normally a single launch): a request is
      // traced against the scene its own camera names.
      for (size_t i = 0, r0 = 0; i < rays.size() && !rc; r0 += rays[i].req.n_rays, ++i) {
        size_t j = i, n_run = rays[i].req.n_rays;
        while (j + 1 < rays.size() && same_scene(rays[j + 1].req.camera, rays[i].req.camera))
          n_run += rays[++j].req.n_rays;
        if (n_run) rc = vx_query_rays(s->ctx, &rays[i].req.camera, d_rays + r0, (uint32_t)n_run, d_hits + r0, s->qstream);
        for (; i < j; ++i) r0 += rays[i].req.n_rays;
      }
      // Hit downloads queue behind the query on its stream (only the ranges of
      // requests that asked for their hits), overlapping the view passes.
      at = 0;
      for (Pending& p : rays) {
        const size_t n = p.req.n_rays;
        vx_hit* dst = p.pin != kNoPin ? s->pin_hits + p.pin : h_hits + at;
        if (!rc && p.req.out_hits && n &&
            hipMemcpyAsync(dst, d_hits + at, n * sizeof(vx_hit), hipMemcpyDeviceToHost, s->qstream) != hipSuccess)
          rc = vx_fail(VX_E_DEVICE, "server: hit download");
        at += n;
      }
    }
    // Views: bind each request to its stream's view (created on its first
    // frame) and write the camera.
    std::vector<vx_view*> vs;
    bool created_view = false;  // the launch allocated a view: not a sample for the cost model
    for (Pending& p : views) {
      Stream* stp = nullptr;
      {
        std::lock_guard<std::mutex> lk(s->mu);
        stp = &s->streams[p.req.stream_id];
      }
      if (!stp->view && !rc) {
        rc = vx_view_create(s->ctx, p.req.width, p.req.height, p.req.width, p.req.height, &stp->view);
        created_view = true;
      }
      if (!rc) rc = vx_write_camera(stp->view, &p.req.camera);
      vs.push_back(stp->view);
    }
    bool rendered = false;
    if (!rc && !vs.empty()) {
      if (!batch) {
        rc = vx_batch_create(s->ctx, vs.data(), (uint32_t)vs.size(), &batch);
        vx_batch_config bc{6u, 1u, 0u, 0u};  // the view set changes per launch: direct launches
        if (!rc) rc = vx_batch_set_config(batch, &bc);
      } else {
        rc = vx_internal_batch_rebind(batch, vs.data(), (uint32_t)vs.size());
      }
      if (!rc) rc = vx_batch_render(batch, VX_PASS_FRAME, nullptr);
      rendered = true;
    }
    // Ray batches complete first: their hits are copied out to the callers
    // while the views still render.
    if (!rays.empty()) {
      int rc_rays = rc;
      if (!rc_rays && nr && hipStreamSynchronize(s->qstream) != hipSuccess)
        rc_rays = vx_fail(VX_E_DEVICE, "server: query stream");
      size_t at = 0;
      for (Pending& p : rays) {
        const size_t n = p.req.n_rays;
        if (!rc_rays && p.req.out_hits && n)
          std::memcpy(p.req.out_hits, p.pin != kNoPin ? s->pin_hits + p.pin : h_hits + at, n * sizeof(vx_hit));
        at += n;
      }
      const double tr = now_ms();
      {
        std::lock_guard<std::mutex> lk(s->mu);
        if (rc_rays) s->last_error = rc_rays;
        for (Pending& p : rays) complete(s, p, rc_rays, t0, tr, s->lat_rays);
      }
      s->cv.notify_all();
      if (!rc) rc = rc_rays;
    }
    if (!rc && rendered) rc = vx_sync(s->ctx);
    for (size_t i = 0; i < views.size() && !rc; ++i)
      if (views[i].req.out_rgba)
        rc = vx_read_output(vs[i], VX_OUT_OUTPUT, views[i].req.out_rgba,
                            (size_t)views[i].req.width * views[i].req.height * 16);
    const double t1 = now_ms();
    {
      std::lock_guard<std::mutex> lk(s->mu);
      if (rc) s->last_error = rc;
      ++s->st.launches;
      s->st.busy_ms += t1 - t0;
      if (!rc && !created_view && !s->inflight.empty() && t1 > t0) s->cost.add(s->inflight, t1 - t0);
      s->inflight = Work{};
      s->last_launch_end_ms = t1;
      for (Pending& p : views) {
        bool more = false;
        for (auto& q : s->inbox) more |= q.req.kind == VX_REQ_VIEW && q.req.stream_id == p.req.stream_id;
        s->streams[p.req.stream_id].busy = more;
        complete(s, p, rc, t0, t1, s->lat_view);
      }
      if (s->calib.size() >= kCalibMin) {
        std::vector<double> r = s->calib;
        const size_t k = (size_t)(0.99 * (double)(r.size() - 1));
        std::nth_element(r.begin(), r.begin() + k, r.end());
        s->margin = std::min(std::max(r[k], 1.0), kMarginMax);
      }
    }
    s->cv.notify_all();
  }
  if (batch) vx_batch_destroy(batch);
  if (d_rays) hipFree(d_rays);
  if (d_hits) hipFree(d_hits);
  if (h_rays) hipHostFree(h_rays);
  if (h_hits) hipHostFree(h_hits);
}

}  // namespace

extern "C" {

int vx_server_create(vx_ctx* ctx, const vx_host_world_view* world, const vx_server_config* cfg, vx_server** out) {
  if (!ctx || !world || !cfg || !out) return vx_fail(VX_E_INVALID, "vx_server_create: null argument");
  *out = nullptr;
  if (world->chunk_count == 0 || !world->meta) return vx_fail(VX_E_INVALID, "vx_server_create: empty world");
  vx_server* s = new (std::nothrow) vx_server();
  if (!s) return vx_fail(VX_E_NOMEM, "vx_server_create: host allocation");
  s->ctx = ctx;
  vx_internal_child(ctx, +1);  // released by vx_server_destroy
  s->cfg = *cfg;
  if (!s->cfg.max_views_per_launch) s->cfg.max_views_per_launch = 64;
  if (!s->cfg.edit_threads) s->cfg.edit_threads = 8;
  s->meta.assign(world->meta, world->meta + world->chunk_count);
  s->origin.resize(world->chunk_count);
  for (uint32_t i = 0; i < world->chunk_count; ++i)
    s->origin[i] = {world->meta[i].origin[0], world->meta[i].origin[1], world->meta[i].origin[2]};
  s->dirty.assign(world->chunk_count, 0);
  s->gen.assign(world->chunk_count, 0);
  s->applied.assign(world->chunk_count, 0);
  std::vector<std::pair<uint64_t, uint64_t>> used;
  for (const vx_chunk_meta& m : s->meta) used.emplace_back(m.node_base, m.node_count);
  s->arena.init(vx_internal_node_capacity(ctx), used);  // the node arena is the context's node capacity
  if (hipSetDevice(vx_internal_device(ctx)) != hipSuccess ||
      hipStreamCreateWithFlags(&s->qstream, hipStreamNonBlocking) != hipSuccess) {
    vx_internal_child(ctx, -1);
    delete s;
    return vx_fail(VX_E_DEVICE, "vx_server_create: query stream");
  }
  // The ray pool: the budget's worth of rays, between 64 Ki and 8 Mi.
  const uint64_t pool = std::min<uint64_t>(std::max<uint64_t>(s->cfg.memory_budget / kRayBytes, 1ull << 16), 1ull << 23);
  if (hipHostMalloc((void**)&s->pin_rays, pool * sizeof(vx_ray), hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&s->pin_hits, pool * sizeof(vx_hit), hipHostMallocDefault) != hipSuccess) {
    vx_server_destroy(s);
    return vx_fail(VX_E_NOMEM, "vx_server_create: pinned ray pool (%llu rays)", (unsigned long long)pool);
  }
  s->pin.init(pool, {});
  s->device_build = (s->cfg.flags & VX_SERVER_DEVICE_BUILD) != 0;
  if (s->device_build) {
    const size_t slots = s->meta.size();
    s->build_batch = (uint32_t)std::min<size_t>(256, slots);
    int rc = VX_OK;
    if (hipMalloc((void**)&s->dmat, slots * 64ull * 64ull * 64ull) != hipSuccess ||
        hipMalloc((void**)&s->dground, slots * 4096ull * sizeof(int32_t)) != hipSuccess)
      rc = vx_fail(VX_E_NOMEM, "vx_server_create: resident material (%zu slots)", slots);
    if (!rc) rc = vx_chunk_builder_create(ctx, s->build_batch, &s->builder);
    if (!rc) rc = fill_device_world(s, s->cfg.edit_threads);
    if (rc) {
      vx_server_destroy(s);
      return rc;
    }
  }
  s->t_start = now_ms();
  if (!s->device_build)
    for (uint32_t i = 0; i < s->cfg.edit_threads; ++i) s->workers.emplace_back(worker_main, s);
  s->device = std::thread(device_main, s);
  *out = s;
  return VX_OK;
}

void vx_server_destroy(vx_server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv.notify_all();
  s->wcv.notify_all();
  if (s->device.joinable()) s->device.join();
  for (auto& w : s->workers)
    if (w.joinable()) w.join();
  for (auto& kv : s->streams)
    if (kv.second.view) vx_view_destroy(kv.second.view);
  if (s->builder) vx_chunk_builder_destroy(s->builder);
  for (void* p : {(void*)s->dmat, (void*)s->dground, (void*)s->d_off, (void*)s->d_val})
    if (p) hipFree(p);
  if (s->qstream) hipStreamDestroy(s->qstream);
  if (s->pin_rays) hipHostFree(s->pin_rays);
  if (s->pin_hits) hipHostFree(s->pin_hits);
  vx_internal_child(s->ctx, -1);
  delete s;
}

int vx_server_submit(vx_server* s, const vx_request* req) {
  if (!s || !req) return vx_fail(VX_E_INVALID, "vx_server_submit: null argument");
  Pending p;
  p.req = *req;
  if (req->kind == VX_REQ_VIEW) {
    if (!req->width || !req->height || (uint64_t)req->width * req->height > (1ull << 26))
      return vx_fail(VX_E_INVALID, "vx_server_submit: bad view size");
  } else if (req->kind == VX_REQ_RAYS) {
    if (!req->rays && req->n_rays) return vx_fail(VX_E_INVALID, "vx_server_submit: rays missing");
    if (req->n_rays) {
      bool pinned = false;
      {
        std::lock_guard<std::mutex> lk(s->mu);
        pinned = s->pin.alloc(req->n_rays, p.pin);
      }
      if (pinned)
        std::memcpy(s->pin_rays + p.pin, req->rays, (size_t)req->n_rays * sizeof(vx_ray));
      else
        p.rays.assign(req->rays, req->rays + req->n_rays);
    }
  } else {
    return vx_fail(VX_E_INVALID, "vx_server_submit: bad request kind %u", req->kind);
  }
  p.submit_ms = now_ms();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    ++s->st.submitted;
    auto reject = [&] {
      ++s->st.rejected;
      if (req->kind == VX_REQ_VIEW && !s->streams[req->stream_id].w) s->streams.erase(req->stream_id);
      if (p.pin != kNoPin) s->pin.release(p.pin, req->n_rays);
    };
    uint64_t need = 0;
    if (req->kind == VX_REQ_VIEW) {
      Stream& stp = s->streams[req->stream_id];
      if (s->cfg.max_queued_frames && stp.busy) {
        uint32_t q = 0;
        for (auto& p2 : s->inbox) q += p2.req.kind == VX_REQ_VIEW && p2.req.stream_id == req->stream_id;
        if (q + 1u > s->cfg.max_queued_frames) {  // counts the frame in flight too
          reject();
          return vx_fail(VX_E_BUSY, "vx_server_submit: stream %u already has %u frame(s) pending", req->stream_id, q + 1u);
        }
      }
      if (stp.w && (stp.w != req->width || stp.h != req->height)) {
        reject();
        return vx_fail(VX_E_INVALID, "vx_server_submit: stream %u is %ux%u", req->stream_id, stp.w, stp.h);
      }
      if (!stp.w) need = (uint64_t)req->width * req->height * kViewBytesPerPx;  // a new stream's resident view
    } else {
      need = (uint64_t)req->n_rays * kRayBytes;  // in-flight ray buffers
    }
    if (s->mem_in_use + need > s->cfg.memory_budget) {
      reject();
      return vx_fail(VX_E_BUSY, "vx_server_submit: %llu bytes over the %llu-byte budget",
                     (unsigned long long)(s->mem_in_use + need), (unsigned long long)s->cfg.memory_budget);
    }
    const Work w = request_work(p);
    // A server that has been idle for kProbeAfterMs admits the request
    // whatever the estimate says: it runs alone, and its launch refreshes the
    // cost model (which otherwise could stay pessimistic for good once every
    // request is rejected).
    const bool probe = s->inbox.empty() && s->inflight.empty() && p.submit_ms - s->last_launch_end_ms > kProbeAfterMs;
    if (s->cfg.latency_target_ms > 0.0f && s->cost.ready() && !probe) {
      // Estimated finish: the rest of the launch in flight, then the launches
      // the queue and this request need (each at most the launch cap), each
      // with its fixed cost.
      const double left =
          s->inflight.empty() ? 0.0 : std::max(0.0, s->cost.ms(s->inflight) - (p.submit_ms - s->inflight_t0));
      Work ahead = s->queued;
      ahead += w;
      const double work_ms = s->cost.rays_ms(ahead);
      const double per_launch = std::max(1e-3, kLaunchShare * s->cfg.latency_target_ms - s->cost.fixed());
      const double launches = std::max(1.0, std::ceil(work_ms / per_launch));
      p.est_ms = left + launches * s->cost.fixed() + work_ms;
      const double est = s->margin * p.est_ms;
      if (est > s->cfg.latency_target_ms) {
        reject();
        return vx_fail(VX_E_BUSY, "vx_server_submit: estimated %.1f ms over the %.1f ms latency target", est,
                       (double)s->cfg.latency_target_ms);
      }
    }
    s->queued += w;
    s->mem_in_use += need;
    if (req->kind == VX_REQ_VIEW) {
      Stream& stp = s->streams[req->stream_id];
      if (!stp.w) {
        stp.w = req->width;
        stp.h = req->height;
        stp.bytes = need;  // released when the stream closes
      }
      stp.busy = true;
    } else {
      p.reserved = need;
    }
    ++s->st.admitted;
    s->inbox.push_back(std::move(p));
  }
  s->cv.notify_all();
  return VX_OK;
}

int vx_server_poll(vx_server* s, vx_result* out) {
  if (!s || !out) return vx_fail(VX_E_INVALID, "vx_server_poll: null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->outbox.empty()) return VX_E_AGAIN;
  *out = s->outbox.front();
  s->outbox.pop_front();
  return VX_OK;
}

int vx_server_edit(vx_server* s, const int32_t* xyz, const uint8_t* mats, uint32_t n) {
  if (!s || (n && (!xyz || !mats))) return vx_fail(VX_E_INVALID, "vx_server_edit: null argument");
  {
    std::lock_guard<std::mutex> lk(s->mu);
    for (uint32_t i = 0; i < n; ++i) {
      uint32_t idx = 0;
      const uint32_t slot = slot_of_voxel(s, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], idx);
      if (slot == VX_INVALID_U32) continue;
      ++s->gen[slot];
      if (s->device_build) {
        s->pend[slot][idx] = mats[i];
        if (!s->dirty[slot]) {
          s->dirty[slot] = 1;
          s->dev_q.push_back(slot);
        }
        continue;
      }
      s->edits[slot][idx] = mats[i];
      if (!s->dirty[slot]) {
        s->dirty[slot] = 1;
        s->rebuild_q.push_back(slot);
      }
    }
  }
  s->wcv.notify_all();
  s->cv.notify_all();
  return VX_OK;
}

int vx_server_close_stream(vx_server* s, uint32_t stream_id) {
  if (!s) return vx_fail(VX_E_INVALID, "vx_server_close_stream: null server");
  vx_view* v = nullptr;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->streams.find(stream_id);
    if (it == s->streams.end()) return vx_fail(VX_E_INVALID, "vx_server_close_stream: no stream %u", stream_id);
    // Wait until no frame of the stream is queued or in flight.  The wait
    // releases the lock: submits may rehash the map and another close may
    // erase the stream meanwhile, so look it up again afterwards.
    s->cv.wait(lk, [&] {
      for (auto& p : s->inbox)
        if (p.req.kind == VX_REQ_VIEW && p.req.stream_id == stream_id) return false;
      const auto cur = s->streams.find(stream_id);
      return cur == s->streams.end() || !cur->second.busy;
    });
    it = s->streams.find(stream_id);
    if (it == s->streams.end()) return vx_fail(VX_E_INVALID, "vx_server_close_stream: stream %u closed meanwhile", stream_id);
    v = it->second.view;
    s->mem_in_use -= it->second.bytes;
    s->streams.erase(it);
  }
  if (v) vx_view_destroy(v);
  return VX_OK;
}

int vx_server_drain(vx_server* s) {
  if (!s) return vx_fail(VX_E_INVALID, "vx_server_drain: null server");
  std::unique_lock<std::mutex> lk(s->mu);
  s->cv.wait(lk, [&] { return s->last_error != VX_OK || (s->inbox.empty() && s->st.completed == s->st.admitted); });
  return s->last_error;
}

int vx_server_wait_edits(vx_server* s) {
  if (!s) return vx_fail(VX_E_INVALID, "vx_server_wait_edits: null server");
  std::unique_lock<std::mutex> lk(s->mu);
  // Every queued edit has been built and applied at a launch boundary, or an
  // error stopped one of them (a failed fill, build or upload), or only
  // builds the node arena cannot place are left: nothing else is queued or
  // running, so no boundary will free room for them.
  bool stuck = false;
  s->cv.wait(lk, [&] {
    if (s->last_error != VX_OK) return true;
    if (!s->rebuild_q.empty() || !s->ready.empty() || s->building || !s->dev_q.empty() || s->dev_busy) return false;
    if (!s->deferred.empty() || !s->dev_deferred.empty()) {
      if (!s->inbox.empty()) return false;  // the next launch boundary retries them
      stuck = true;
      return true;
    }
    for (size_t i = 0; i < s->gen.size(); ++i)
      if (s->applied[i] < s->gen[i]) return false;
    return true;
  });
  if (s->last_error == VX_OK && stuck)
    return vx_fail(VX_E_CAPACITY, "vx_server_wait_edits: rebuilt chunks do not fit the node arena");
  return s->last_error;
}

int vx_server_get_stats(vx_server* s, vx_server_stats* out) {
  if (!s || !out) return vx_fail(VX_E_INVALID, "vx_server_get_stats: null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  *out = s->st;
  out->view_p50_ms = pct(s->lat_view, 50.0);
  out->view_p99_ms = pct(s->lat_view, 99.0);
  out->rays_p50_ms = pct(s->lat_rays, 50.0);
  out->rays_p99_ms = pct(s->lat_rays, 99.0);
  out->wall_ms = now_ms() - s->t_start;
  out->memory_in_use = s->mem_in_use;
  return VX_OK;
}

}  // extern "C"
