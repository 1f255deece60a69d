// vx_builder.hip — the chunk builder on the device: the same arenas as
// vx_build_chunk (svo/builder.rs:742-986, builder_svo.rs:69-534,
// builder_ropes.rs:31-195), byte for byte, for many chunks per launch.
//
// One workgroup (1024 lanes) per 64^3 chunk, in two launches:
//
//   count  Bottom-up over the octree levels (5: 2^3 cells of voxels ... 0: the
//          root), one cell per lane-iteration: the child-existence mask, whether
//          the cell is full, and its kind: absent, leaf (a voxel; a full cube
//          entirely below ground - dirt_depth collapsed to STONE; or eight leaf
//          children of one material merged) or interior.  Then, per level, the
//          exclusive scan of interior cells' child counts in the builder's cell
//          order (y-slab, i = (y*side + z)*side + x): the index of every cell's
//          first child within the next level's packed node list.  The level
//          totals give the chunk's node count.
//   emit   Top-down, level by level: each interior cell writes its children's
//          nodes at base[level] + scan + rank and their ropes (sibling inside
//          the parent, else the parent's rope descended one level), then the
//          colinfo words (top solid voxel per column) and the macro words
//          (occupancy of the 8^3-voxel cells, which are the level-3 cells).
//
// Between the launches the host places each chunk (node_base, macro_base,
// colinfo_base), so nodes and ropes are written straight into the context's
// resident arenas.  Per-chunk state lives in a workspace in HBM (about 450 KB
// per chunk: three u32 per cell of levels 0-5); a level is complete before the
// next one reads it (workgroup barrier).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "vx_internal.h"

namespace {

constexpr uint32_t kInvalid = 0xFFFFFFFFu;
constexpr uint32_t kThreads = 1024;
constexpr uint32_t kVoxels = 64u * 64u * 64u;
// First cell of each level 0..5 in a chunk's per-level arrays (side 2^L).
__host__ __device__ constexpr uint32_t level_first(uint32_t L) { return L == 0 ? 0u : level_first(L - 1) + (1u << (3 * (L - 1))); }
constexpr uint32_t kCells = level_first(6);  // 37449
// Workspace per chunk (u32): info, scan and node index per cell, 16 words of
// chunk state (base[0..7], base[7] = node count), then the 2048 colinfo words.
constexpr uint32_t kWsInfo = 0, kWsScan = kCells, kWsIdx = 2 * kCells, kWsMeta = 3 * kCells;
constexpr uint32_t kWsCol = (3 * kCells + 16 + 63) & ~63u;
constexpr uint32_t kWsStride = kWsCol + 2048u;

// Cell info word.
constexpr uint32_t kExists = 1u, kLeaf = 2u, kFull = 4u;
constexpr uint32_t kVisited = 0x80000000u;  // node index word: the rope pass reached this cell

__device__ inline uint32_t info_mat(uint32_t i) { return (i >> 8) & 0xFFu; }
__device__ inline uint32_t info_mask(uint32_t i) { return (i >> 16) & 0xFFu; }
__device__ inline bool interior(uint32_t i) { return (i & (kExists | kLeaf)) == kExists; }

__device__ inline uint32_t cell_index(uint32_t side_bits, uint32_t x, uint32_t y, uint32_t z) {
  return (((y << side_bits) + z) << side_bits) + x;
}

// Octree key (svo/builder_svo.rs:22-35) of the level-L cell (x, y, z).
__device__ inline uint32_t cell_key(uint32_t L, uint32_t x, uint32_t y, uint32_t z) {
  return (x & 63u) | ((y & 63u) << 6) | ((z & 63u) << 12) | ((L & 7u) << 18);
}

// Info of a level-6 cell (voxel).
__device__ inline uint32_t voxel_info(uint32_t m) { return m ? (kExists | kLeaf | kFull | (m << 8)) : 0u; }

// Block-wide exclusive scan of one value per lane (1024 lanes, 16 wavefronts).
__device__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* sh16, uint32_t& total) {
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  uint32_t x = v;
  for (uint32_t d = 1; d < 64; d <<= 1) {
    const uint32_t y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  if (lane == 63) sh16[wave] = x;
  __syncthreads();
  if (wave == 0) {
    uint32_t w = lane < 16 ? sh16[lane] : 0u;
    for (uint32_t d = 1; d < 16; d <<= 1) {
      const uint32_t y = __shfl_up(w, d, 64);
      if (lane >= d) w += y;
    }
    if (lane < 16) sh16[lane] = w;
  }
  __syncthreads();
  const uint32_t before = wave ? sh16[wave - 1] : 0u;
  total = sh16[15];
  __syncthreads();
  return before + x - v;
}

// A parent cell's info from its eight children's (ci = hx | hy<<1 | hz<<2),
// builder_svo.rs:95-255: absent if no child exists; STONE leaf if the full cube
// lies entirely below min ground - dirt_depth; one leaf if all eight children
// are leaves of one material; else interior.
__device__ inline uint32_t parent_info(const uint32_t ch[8], bool below_stone_line) {
  uint32_t mask = 0, full = 1, leaves = 1;
  const uint32_t m0 = info_mat(ch[0]);
  bool same = true;
  for (uint32_t ci = 0; ci < 8; ++ci) {
    const uint32_t c = ch[ci];
    if (c & kExists) mask |= 1u << ci;
    full &= (c >> 2) & 1u;
    leaves &= (c >> 1) & 1u;
    same = same && info_mat(c) == m0;
  }
  if (!mask) return 0u;
  const uint32_t f = full ? kFull : 0u;
  if (below_stone_line && full) return kExists | kLeaf | f | (VX_MAT_STONE << 8) | (mask << 16);
  if (mask == 0xFFu && leaves && same) return kExists | kLeaf | f | (m0 << 8) | (mask << 16);
  return kExists | f | (mask << 16);
}

// First entry of the ground-minimum level for cubes of 2^k voxels (k = 1..6,
// grids of 32^2 ... 1^2 columns) in the packed 1365-entry array.
__device__ inline uint32_t gmip_first(uint32_t k) {
  uint32_t f = 0;
  for (uint32_t j = 1; j < k; ++j) f += (64u >> j) * (64u >> j);
  return f;
}

struct CountArgs {
  const uint8_t* material;  // chunks of 64^3, y-slab
  const int32_t* ground;    // chunks of 64^2
  const uint32_t* index;    // n: batch chunk -> position in material/ground (null: identity)
  const int32_t* chunk_oy;  // n
  int32_t dirt_depth;
  uint32_t* ws;
  uint32_t* counts;         // 2n: node count, holds a solid voxel
};

__global__ __launch_bounds__(kThreads) void k_build_count(CountArgs a) {
  const uint32_t chunk = blockIdx.x, t = threadIdx.x;
  const uint32_t src = a.index ? a.index[chunk] : chunk;
  const uint8_t* mat = a.material + (size_t)src * kVoxels;
  const int32_t* gr = a.ground + (size_t)src * 4096u;
  uint32_t* ws = a.ws + (size_t)chunk * kWsStride;
  uint32_t* info = ws + kWsInfo;
  const int32_t oy = a.chunk_oy[chunk];

  // Ground minimum mip (svo/mips.rs:11-113), levels of cube size 2..64.
  __shared__ int32_t gmin[1365];
  __shared__ uint32_t sh16[16];
  __shared__ uint32_t top[4096];  // per column (z*64 + x): ((y+1) << 8 | material) of its highest solid voxel, 0 if none
  for (uint32_t c = t; c < 4096u; c += kThreads) top[c] = 0u;
  {
    const uint32_t x = t & 31u, z = t >> 5;
    const int32_t* r0 = gr + (2 * z) * 64 + 2 * x;
    gmin[t] = min(min(r0[0], r0[1]), min(r0[64], r0[65]));
  }
  __syncthreads();
  for (uint32_t s = 2, side = 16; s <= 6; ++s, side >>= 1) {
    if (t < side * side) {
      const uint32_t x = t % side, z = t / side, ps = side * 2;
      const int32_t* p = gmin + gmip_first(s - 1) + (2 * z) * ps + 2 * x;
      gmin[gmip_first(s) + t] = min(min(p[0], p[1]), min(p[ps], p[ps + 1]));
    }
    __syncthreads();
  }

  // Node index words start "not reached": an interior cell under a collapsed
  // STONE leaf is never emitted itself, yet its children are.
  for (uint32_t c = t; c < kCells; c += kThreads) ws[kWsIdx + c] = 0u;

  // Levels 5..0, bottom-up.
  for (int L = 5; L >= 0; --L) {
    const uint32_t bits = (uint32_t)L, side = 1u << bits, ncell = side * side * side;
    const uint32_t size = 64u >> bits;  // voxels per cell edge
    uint32_t* out = info + level_first(L);
    const uint32_t* child = info + level_first(L + 1 < 6 ? L + 1 : 5);
    for (uint32_t c = t; c < ncell; c += kThreads) {
      const uint32_t x = c & (side - 1), z = (c >> bits) & (side - 1), y = c >> (2 * bits);
      uint32_t ch[8];
      if (L == 5) {
        for (uint32_t ci = 0; ci < 8; ++ci) {
          const uint32_t vx = 2 * x + (ci & 1u), vy = 2 * y + ((ci >> 1) & 1u), vz = 2 * z + (ci >> 2);
          ch[ci] = voxel_info(mat[(vy << 12) + (vz << 6) + vx]);
        }
        // Column tops: the upper voxel of each of the cell's four columns wins over the lower.
        for (uint32_t q = 0; q < 4; ++q) {
          const uint32_t lo = ch[(q & 1u) | ((q >> 1) << 2)], hi = ch[(q & 1u) | 2u | ((q >> 1) << 2)];
          const uint32_t e = hi ? ((2 * y + 2) << 8) | info_mat(hi) : lo ? ((2 * y + 1) << 8) | info_mat(lo) : 0u;
          if (e) atomicMax(&top[((2 * z + (q >> 1)) << 6) + 2 * x + (q & 1u)], e);
        }
      } else {
        for (uint32_t ci = 0; ci < 8; ++ci)
          ch[ci] = child[cell_index(bits + 1, 2 * x + (ci & 1u), 2 * y + ((ci >> 1) & 1u), 2 * z + (ci >> 2))];
      }
      const int32_t g = gmin[gmip_first(6 - bits) + z * side + x];
      const int32_t y1 = oy + (int32_t)(y * size) + (int32_t)size - 1;
      out[c] = parent_info(ch, y1 < g - a.dirt_depth);
    }
    __syncthreads();
  }

  // Colinfo words (svo/builder.rs:742-805): per column the highest solid
  // voxel's y and material (y = 255, material 0 for an empty column), two
  // columns per word, low half first.
  for (uint32_t w = t; w < 2048u; w += kThreads) {
    uint32_t word = 0;
    for (uint32_t h = 0; h < 2; ++h) {
      const uint32_t e = top[2 * w + h];
      word |= (e ? ((e >> 8) - 1u) | ((e & 0xFFu) << 8) : 255u) << (16 * h);
    }
    ws[kWsCol + w] = word;
  }

  // Per level, the scan of interior cells' child counts: first child's index in
  // the next level's packed list.  base[L+1] = base[L] + level L's total.
  uint32_t base = 1;  // node 0 is the root
  if (t == 0) {
    ws[kWsMeta + 0] = 0;
    ws[kWsMeta + 1] = 1;
  }
  for (uint32_t L = 0; L <= 5; ++L) {
    const uint32_t ncell = 1u << (3 * L);
    const uint32_t per = (ncell + kThreads - 1) / kThreads;
    const uint32_t c0 = min(t * per, ncell), c1 = min(c0 + per, ncell);
    const uint32_t* in = info + level_first(L);
    uint32_t* scan = ws + kWsScan + level_first(L);
    uint32_t sum = 0;
    for (uint32_t c = c0; c < c1; ++c) {
      const uint32_t i = in[c];
      sum += interior(i) ? (uint32_t)__popc(info_mask(i)) : 0u;
    }
    uint32_t total;
    uint32_t run = block_exclusive_scan(sum, sh16, total);
    for (uint32_t c = c0; c < c1; ++c) {
      scan[c] = run;
      const uint32_t i = in[c];
      run += interior(i) ? (uint32_t)__popc(info_mask(i)) : 0u;
    }
    base += total;
    if (t == 0) ws[kWsMeta + L + 2] = base;
  }
  if (t == 0) {
    a.counts[2 * chunk] = base;
    a.counts[2 * chunk + 1] = info[0] & kExists;
  }
}

struct EmitArgs {
  const uint8_t* material;
  const uint32_t* index;
  uint32_t* ws;
  const uint64_t* node_base;     // n
  const uint32_t* macro_base;    // n
  const uint32_t* colinfo_base;  // n
  vx_node* nodes;
  vx_node_ropes* ropes;
  uint32_t* macro;
  uint32_t* colinfo;
  uint32_t* err;  // set when a chunk's workspace is inconsistent
};

// Rope target one level down (builder_ropes.rs:31-75): an invalid rope stays
// invalid, a leaf neighbour is the target itself, an interior neighbour gives
// its child on the near side (or invalid when that child is absent).  The
// neighbour's child_base and child_mask are passed in: a parent reads its six
// rope targets' nodes once for all of its children.
__device__ inline uint32_t descend_one(uint32_t nei, uint32_t nbase, uint32_t nmask, uint32_t ci) {
  if (nei == kInvalid) return kInvalid;
  if (nbase == kInvalid) return nei;
  const uint32_t bit = 1u << ci;
  if (!(nmask & bit)) return kInvalid;
  return nbase + (uint32_t)__popc(nmask & (bit - 1u));
}

__global__ __launch_bounds__(kThreads) void k_build_emit(EmitArgs a) {
  const uint32_t chunk = blockIdx.x, t = threadIdx.x;
  const uint8_t* mat = a.material + (size_t)(a.index ? a.index[chunk] : chunk) * kVoxels;
  uint32_t* ws = a.ws + (size_t)chunk * kWsStride;
  const uint32_t* info = ws + kWsInfo;
  const uint32_t* scan = ws + kWsScan;
  uint32_t* idx = ws + kWsIdx;
  vx_node* nodes = a.nodes + a.node_base[chunk];
  vx_node_ropes* ropes = a.ropes + a.node_base[chunk];
  const uint32_t total = ws[kWsMeta + 7];

  // The root: its own node and all-invalid ropes.
  if (t == 0) {
    const uint32_t r = info[0];
    vx_node n;
    if (!(r & kExists)) {
      n = {kInvalid, 0u, VX_MAT_AIR, cell_key(0, 0, 0, 0)};
    } else if (r & kLeaf) {
      n = {kInvalid, 0u, info_mat(r), cell_key(0, 0, 0, 0)};
    } else {
      n = {ws[kWsMeta + 1] + scan[0], info_mask(r), 0u, cell_key(0, 0, 0, 0)};
    }
    nodes[0] = n;
    ropes[0] = {kInvalid, kInvalid, kInvalid, kInvalid, kInvalid, kInvalid, 0u, 0u};
    idx[0] = 0u | kVisited;
  }
  __syncthreads();

  // Levels 1..6: every interior cell of level L-1 writes its children.
  for (uint32_t L = 1; L <= 6; ++L) {
    const uint32_t pbits = L - 1, pside = 1u << pbits, ncell = pside * pside * pside;
    const uint32_t* pinfo = info + level_first(L - 1);
    const uint32_t* pscan = scan + level_first(L - 1);
    const uint32_t* pidx = idx + level_first(L - 1);
    const uint32_t child_base_level = L < 6 ? ws[kWsMeta + L + 1] : 0u;  // base[L+1]
    const uint32_t first = ws[kWsMeta + L];                             // base[L]
    for (uint32_t p = t; p < ncell; p += kThreads) {
      const uint32_t pi = pinfo[p];
      if (!interior(pi)) continue;
      const uint32_t px = p & (pside - 1), pz = (p >> pbits) & (pside - 1), py = p >> (2 * pbits);
      const uint32_t pmask = info_mask(pi);
      const uint32_t pw = pidx[p];
      const bool visited = (pw & kVisited) != 0;
      vx_node_ropes pr = {kInvalid, kInvalid, kInvalid, kInvalid, kInvalid, kInvalid, 0u, 0u};
      if (visited) pr = ropes[pw & ~kVisited];
      // The rope targets are nodes of earlier levels (complete behind the level
      // barrier): their words are loaded together, not once per child between
      // the children's stores.
      const uint32_t nei[6] = {pr.px, pr.nx, pr.py, pr.ny, pr.pz, pr.nz};
      uint32_t nbase[6], nmask[6];
      for (uint32_t d = 0; d < 6; ++d) {
        nbase[d] = kInvalid;
        nmask[d] = 0u;
        if (nei[d] != kInvalid) {
          const vx_node nn = nodes[nei[d]];
          nbase[d] = nn.child_base;
          nmask[d] = nn.child_mask;
        }
      }
      uint32_t kid[8];
      uint32_t rank = 0;
      for (uint32_t ci = 0; ci < 8; ++ci) kid[ci] = (pmask >> ci) & 1u ? first + pscan[p] + rank++ : kInvalid;
      for (uint32_t ci = 0; ci < 8; ++ci) {
        if (kid[ci] == kInvalid) continue;
        if (kid[ci] >= total) {  // cannot happen for a consistent workspace; never write past the chunk
          atomicOr(a.err, 1u);
          continue;
        }
        const uint32_t hx = ci & 1u, hy = (ci >> 1) & 1u, hz = ci >> 2;
        const uint32_t cx = 2 * px + hx, cy = 2 * py + hy, cz = 2 * pz + hz;
        const uint32_t c = cell_index(L, cx, cy, cz);
        const uint32_t ci_info = L == 6 ? voxel_info(mat[(cy << 12) + (cz << 6) + cx]) : info[level_first(L) + c];
        vx_node n;
        if (ci_info & kLeaf) {
          n = {kInvalid, 0u, info_mat(ci_info), cell_key(L, cx, cy, cz)};
        } else {
          n = {child_base_level + scan[level_first(L) + c], info_mask(ci_info), 0u, cell_key(L, cx, cy, cz)};
        }
        nodes[kid[ci]] = n;
        vx_node_ropes r = {kInvalid, kInvalid, kInvalid, kInvalid, kInvalid, kInvalid, 0u, 0u};
        if (visited) {
          const uint32_t sx = kid[ci ^ 1u], sy = kid[ci ^ 2u], sz = kid[ci ^ 4u];
          r.px = (hx == 0 && sx != kInvalid) ? sx : descend_one(nei[0], nbase[0], nmask[0], 0u | (hy << 1) | (hz << 2));
          r.nx = (hx == 1 && sx != kInvalid) ? sx : descend_one(nei[1], nbase[1], nmask[1], 1u | (hy << 1) | (hz << 2));
          r.py = (hy == 0 && sy != kInvalid) ? sy : descend_one(nei[2], nbase[2], nmask[2], hx | (0u << 1) | (hz << 2));
          r.ny = (hy == 1 && sy != kInvalid) ? sy : descend_one(nei[3], nbase[3], nmask[3], hx | (1u << 1) | (hz << 2));
          r.pz = (hz == 0 && sz != kInvalid) ? sz : descend_one(nei[4], nbase[4], nmask[4], hx | (hy << 1) | (0u << 2));
          r.nz = (hz == 1 && sz != kInvalid) ? sz : descend_one(nei[5], nbase[5], nmask[5], hx | (hy << 1) | (1u << 2));
        }
        ropes[kid[ci]] = r;
        if (L < 6) idx[level_first(L) + c] = kid[ci] | (visited ? kVisited : 0u);
      }
    }
    __syncthreads();
  }

  // Colinfo words, computed by the count step.
  uint32_t* col = a.colinfo + a.colinfo_base[chunk];
  for (uint32_t w = t; w < 2048u; w += kThreads) col[w] = ws[kWsCol + w];
  // Macro occupancy (svo/builder.rs:809-841): bit mx + 8*(my + 8*mz) for each
  // 8^3-voxel cell holding a solid voxel, i.e. each present level-3 cell.
  if (t < 16) {
    uint32_t word = 0;
    for (uint32_t b = 0; b < 32; ++b) {
      const uint32_t bit = t * 32 + b;
      const uint32_t mx = bit & 7u, my = (bit >> 3) & 7u, mz = bit >> 6;
      if (info[level_first(3) + cell_index(3, mx, my, mz)] & kExists) word |= 1u << b;
    }
    a.macro[a.macro_base[chunk] + t] = word;
  }
}

// Edits: material[offset[i]] = value[i] (offsets distinct).
__global__ __launch_bounds__(256) void k_scatter_u8(uint8_t* __restrict__ dst, const uint64_t* __restrict__ off,
                                                    const uint8_t* __restrict__ val, uint32_t n) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i < n) dst[off[i]] = val[i];
}

}  // namespace

struct vx_chunk_builder {
  vx_ctx* ctx = nullptr;
  uint32_t max_chunks = 0;
  uint32_t* ws = nullptr;        // max_chunks * kWsStride
  uint32_t* counts = nullptr;    // 2 * max_chunks: node count, solid
  uint32_t* index = nullptr;     // max_chunks (indexed batches)
  int32_t* oy = nullptr;         // max_chunks
  uint64_t* node_base = nullptr; // max_chunks
  uint32_t* macro_base = nullptr;
  uint32_t* colinfo_base = nullptr;
  uint32_t* err = nullptr;
  // The counted batch, pending its emit.
  uint32_t n = 0;
  const uint8_t* material = nullptr;
  bool indexed = false;
  std::vector<uint32_t> host_counts;  // 2 per chunk
};

namespace {

struct Guard {
  int prev = 0;
  explicit Guard(int dev) {
    hipGetDevice(&prev);
    hipSetDevice(dev);
  }
  ~Guard() { hipSetDevice(prev); }
};

#define VXB_HIP(call)                                                                    \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess) return vx_fail(VX_E_DEVICE, "%s: %s", #call, hipGetErrorString(e_)); \
  } while (0)

template <class T>
int balloc(T** p, size_t count) {
  *p = nullptr;
  hipError_t e = hipMalloc(reinterpret_cast<void**>(p), std::max<size_t>(count, 1) * sizeof(T));
  if (e != hipSuccess) return vx_fail(VX_E_NOMEM, "vx_chunk_builder: hipMalloc(%zu bytes): %s", count * sizeof(T),
                                      hipGetErrorString(e));
  return VX_OK;
}

}  // namespace

extern "C" {

int vx_chunk_builder_create(vx_ctx* ctx, uint32_t max_chunks, vx_chunk_builder** out) {
  if (!ctx || !out || max_chunks == 0 || max_chunks > VX_BUILD_CHUNKS_MAX)
    return vx_fail(VX_E_INVALID, "vx_chunk_builder_create: bad argument (max_chunks %u, at most %u)", max_chunks,
                   VX_BUILD_CHUNKS_MAX);
  *out = nullptr;
  vx_internal_arenas ar;
  vx_internal_get_arenas(ctx, &ar);
  vx_chunk_builder* b = new (std::nothrow) vx_chunk_builder();
  if (!b) return vx_fail(VX_E_NOMEM, "vx_chunk_builder_create: host allocation");
  b->ctx = ctx;
  vx_internal_child(ctx, +1);  // released by vx_chunk_builder_destroy, also on the failure path below
  b->max_chunks = max_chunks;
  b->host_counts.assign(2 * (size_t)max_chunks, 0u);
  Guard g(ar.device);
  int rc = VX_OK;
  if ((rc = balloc(&b->ws, (size_t)max_chunks * kWsStride)) || (rc = balloc(&b->counts, 2 * (size_t)max_chunks)) ||
      (rc = balloc(&b->index, max_chunks)) ||
      (rc = balloc(&b->oy, max_chunks)) || (rc = balloc(&b->node_base, max_chunks)) ||
      (rc = balloc(&b->macro_base, max_chunks)) || (rc = balloc(&b->colinfo_base, max_chunks)) ||
      (rc = balloc(&b->err, 1))) {
    vx_chunk_builder_destroy(b);
    return rc;
  }
  *out = b;
  return VX_OK;
}

void vx_chunk_builder_destroy(vx_chunk_builder* b) {
  if (!b) return;
  vx_internal_arenas ar;
  vx_internal_get_arenas(b->ctx, &ar);
  Guard g(ar.device);
  hipStreamSynchronize((hipStream_t)ar.stream);
  for (void* p : {(void*)b->ws, (void*)b->counts, (void*)b->index, (void*)b->oy, (void*)b->node_base, (void*)b->macro_base,
                  (void*)b->colinfo_base, (void*)b->err})
    if (p) hipFree(p);
  vx_internal_child(b->ctx, -1);
  delete b;
}

}  // extern "C"

int vx_internal_builder_count(vx_chunk_builder* b, uint32_t n, const uint8_t* material, const int32_t* ground,
                              const uint32_t* index, const int32_t* chunk_oy, int32_t dirt_depth, uint32_t* node_counts,
                              uint8_t* solid) {
  if (!b || (n && (!material || !ground || !chunk_oy || !node_counts)))
    return vx_fail(VX_E_INVALID, "vx_chunk_builder_count: null argument");
  if (n > b->max_chunks) return vx_fail(VX_E_INVALID, "vx_chunk_builder_count: %u chunks > %u", n, b->max_chunks);
  b->n = 0;
  b->material = nullptr;
  if (!n) return VX_OK;
  vx_internal_arenas ar;
  vx_internal_get_arenas(b->ctx, &ar);
  hipStream_t st = (hipStream_t)ar.stream;
  Guard g(ar.device);
  VXB_HIP(hipMemcpyAsync(b->oy, chunk_oy, n * sizeof(int32_t), hipMemcpyHostToDevice, st));
  if (index) VXB_HIP(hipMemcpyAsync(b->index, index, n * sizeof(uint32_t), hipMemcpyHostToDevice, st));
  CountArgs a{material, ground, index ? b->index : nullptr, b->oy, dirt_depth, b->ws, b->counts};
  hipLaunchKernelGGL(k_build_count, dim3(n), dim3(kThreads), 0, st, a);
  VXB_HIP(hipGetLastError());
  VXB_HIP(hipMemcpyAsync(b->host_counts.data(), b->counts, 2 * n * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  VXB_HIP(hipStreamSynchronize(st));
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t c = b->host_counts[2 * i];
    if (c == 0 || c > VX_MAX_CHUNK_NODES) return vx_fail(VX_E_DEVICE, "vx_chunk_builder_count: chunk %u counted %u nodes", i, c);
    node_counts[i] = c;
    if (solid) solid[i] = (uint8_t)(b->host_counts[2 * i + 1] != 0);
  }
  b->n = n;
  b->material = material;
  b->indexed = index != nullptr;
  return VX_OK;
}

int vx_internal_scatter_u8(vx_ctx* ctx, uint8_t* dst, const uint64_t* d_off, const uint8_t* d_val, uint32_t n) {
  if (!n) return VX_OK;
  vx_internal_arenas ar;
  vx_internal_get_arenas(ctx, &ar);
  Guard g(ar.device);
  hipLaunchKernelGGL(k_scatter_u8, dim3((n + 255u) / 256u), dim3(256), 0, (hipStream_t)ar.stream, dst, d_off, d_val, n);
  VXB_HIP(hipGetLastError());
  return VX_OK;
}

extern "C" {

int vx_chunk_builder_count(vx_chunk_builder* b, uint32_t n, const uint8_t* material, const int32_t* ground,
                           const int32_t* chunk_oy, int32_t dirt_depth, uint32_t* node_counts) {
  return vx_internal_builder_count(b, n, material, ground, nullptr, chunk_oy, dirt_depth, node_counts, nullptr);
}

int vx_chunk_builder_emit(vx_chunk_builder* b, const uint64_t* node_base, const uint32_t* macro_base,
                          const uint32_t* colinfo_base) {
  if (!b || !node_base || !macro_base || !colinfo_base) return vx_fail(VX_E_INVALID, "vx_chunk_builder_emit: null argument");
  if (!b->material) return vx_fail(VX_E_INVALID, "vx_chunk_builder_emit: no counted batch");
  vx_internal_arenas ar;
  vx_internal_get_arenas(b->ctx, &ar);
  const uint32_t n = b->n;
  for (uint32_t i = 0; i < n; ++i) {
    const uint64_t cnt = b->host_counts[2 * i];
    if (node_base[i] > ar.node_cap || cnt > ar.node_cap - node_base[i])
      return vx_fail(VX_E_CAPACITY, "vx_chunk_builder_emit: chunk %u nodes [%llu, +%llu) exceed capacity %llu", i,
                     (unsigned long long)node_base[i], (unsigned long long)cnt, (unsigned long long)ar.node_cap);
    if ((uint64_t)macro_base[i] + 16 > (uint64_t)ar.chunk_cap * 16 ||
        (uint64_t)colinfo_base[i] + 2048 > (uint64_t)ar.chunk_cap * 2048)
      return vx_fail(VX_E_CAPACITY, "vx_chunk_builder_emit: chunk %u macro/colinfo base past capacity", i);
  }
  hipStream_t st = (hipStream_t)ar.stream;
  Guard g(ar.device);
  VXB_HIP(hipMemcpyAsync(b->node_base, node_base, n * sizeof(uint64_t), hipMemcpyHostToDevice, st));
  VXB_HIP(hipMemcpyAsync(b->macro_base, macro_base, n * sizeof(uint32_t), hipMemcpyHostToDevice, st));
  VXB_HIP(hipMemcpyAsync(b->colinfo_base, colinfo_base, n * sizeof(uint32_t), hipMemcpyHostToDevice, st));
  VXB_HIP(hipMemsetAsync(b->err, 0, sizeof(uint32_t), st));
  EmitArgs a{b->material, b->indexed ? b->index : nullptr, b->ws, b->node_base, b->macro_base, b->colinfo_base,
             ar.nodes, ar.ropes, ar.macro, ar.colinfo, b->err};
  hipLaunchKernelGGL(k_build_emit, dim3(n), dim3(kThreads), 0, st, a);
  VXB_HIP(hipGetLastError());
  uint32_t err = 0;
  VXB_HIP(hipMemcpyAsync(&err, b->err, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  VXB_HIP(hipStreamSynchronize(st));
  vx_internal_nodes_changed(b->ctx);
  b->n = 0;
  b->material = nullptr;
  if (err) return vx_fail(VX_E_DEVICE, "vx_chunk_builder_emit: inconsistent workspace");
  return VX_OK;
}

}  // extern "C"
