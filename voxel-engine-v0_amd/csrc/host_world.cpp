// host_world.cpp — CPU producers of the renderer inputs.
//
//   * vx_build_chunk: the reference chunk pipeline after material fill
//     (svo/builder.rs:773-986): column tops, colinfo packing, summed-volume
//     table, macro occupancy bits, ground min/max mip, bottom-up sparse octree
//     and ropes.  Output arrays are byte-identical to the reference's for the
//     same material grid (same node order, same keys, same rope targets).
//   * The seeded synthetic world used by every workload (2-octave value noise,
//     layered terrain, trees).  The reference's own generator
//     (world/generator.rs) is not reproduced.
//   * Host world assembly (slot/meta/grid packing, streaming/manager/slots.rs:
//     231-251 and grid.rs:15-72), the camera packer (app/app.rs:411-511) and the
//     clipmap fill (clipmap.rs:62-122).
#include "voxhip.h"
#include "vx_internal.h"
#include "vx_terrain.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <new>
#include <thread>
#include <vector>

namespace {

constexpr uint32_t kInvalid = 0xFFFFFFFFu;
constexpr int kSide = 64;
constexpr int kSide2 = kSide * kSide;
constexpr int kSide3 = kSide2 * kSide;

// ---------------------------------------------------------------------------
// Octree key (svo/builder_svo.rs:22-35): level 0 = root (size 64), 6 = voxel.
// ---------------------------------------------------------------------------
inline uint32_t ctz32(uint32_t v) { return (uint32_t)__builtin_ctz(v); }

inline uint32_t pack_key(uint32_t chunk_size, int ox, int oy, int oz, int size) {
  uint32_t size_u = (uint32_t)size;
  uint32_t lvl = ctz32(chunk_size) - ctz32(size_u);
  uint32_t sh = ctz32(size_u);
  uint32_t cx = ((uint32_t)ox) >> sh, cy = ((uint32_t)oy) >> sh, cz = ((uint32_t)oz) >> sh;
  return (cx & 63u) | ((cy & 63u) << 6) | ((cz & 63u) << 12) | ((lvl & 7u) << 18);
}

inline vx_node make_leaf(int ox, int oy, int oz, int size, uint32_t mat) {
  vx_node n;
  n.child_base = kInvalid;
  n.child_mask = 0;
  n.material = mat;
  n.key = pack_key(kSide, ox, oy, oz, size);
  return n;
}

// Linear cell index with the y-slab layout used for every dense grid of the
// builder: i = y*side^2 + z*side + x (svo/builder_svo.rs:53-67).
inline int lin_xyz(int side, int x, int y, int z) { return (y * side + z) * side + x; }

// ---------------------------------------------------------------------------
// Summed-volume table, prefix[(z*65 + y)*65 + x] = #solid in [0,x)x[0,y)x[0,z)
// (svo/builder_prefix.rs:9-135).  Values are exact integers, so any summation
// order gives the reference's table.
// ---------------------------------------------------------------------------
struct Prefix {
  static constexpr int D = kSide + 1;
  std::vector<uint32_t> v;
  void build(const uint8_t* mat) {
    v.assign((size_t)D * D * D, 0u);
    for (int z = 1; z <= kSide; ++z)
      for (int y = 1; y <= kSide; ++y) {
        uint32_t run = 0;
        const uint8_t* row = mat + (size_t)(y - 1) * kSide2 + (size_t)(z - 1) * kSide;
        uint32_t* out = &v[((size_t)z * D + y) * D];
        for (int x = 1; x <= kSide; ++x) {
          run += row[x - 1] != VX_MAT_AIR;
          out[x] = run;
        }
      }
    for (int z = 1; z <= kSide; ++z)
      for (int x = 1; x <= kSide; ++x) {
        uint32_t run = 0;
        for (int y = 1; y <= kSide; ++y) {
          size_t i = ((size_t)z * D + y) * D + x;
          run += v[i];
          v[i] = run;
        }
      }
    for (int x = 1; x <= kSide; ++x)
      for (int y = 1; y <= kSide; ++y) {
        size_t base = (size_t)y * D + x;
        uint32_t run = v[base];
        for (int z = 1; z <= kSide; ++z) {
          size_t i = (size_t)z * D * D + base;
          run += v[i];
          v[i] = run;
        }
      }
  }
  inline uint32_t at(int x, int y, int z) const { return v[((size_t)z * D + y) * D + x]; }
  uint32_t cube(int x0, int y0, int z0, int s) const {
    int x1 = x0 + s, y1 = y0 + s, z1 = z0 + s;
    int64_t r = (int64_t)at(x1, y1, z1) - at(x0, y1, z1) - at(x1, y0, z1) - at(x1, y1, z0) +
                at(x0, y0, z1) + at(x0, y1, z0) + at(x1, y0, z0) - at(x0, y0, z0);
    return (uint32_t)r;
  }
};

// Ground min mip (svo/mips.rs:11-113); only the minimum is consulted.
struct GroundMinMip {
  std::vector<int32_t> lv[7];
  void build(const int32_t* ground) {
    lv[0].assign(ground, ground + kSide2);
    int side = kSide;
    for (int l = 1; l < 7; ++l) {
      int ns = side / 2;
      lv[l].resize((size_t)ns * ns);
      for (int z = 0; z < ns; ++z)
        for (int x = 0; x < ns; ++x) {
          const int32_t* a = &lv[l - 1][(size_t)(2 * z) * side + 2 * x];
          const int32_t* b = a + side;
          lv[l][(size_t)z * ns + x] = std::min(std::min(a[0], a[1]), std::min(b[0], b[1]));
        }
      side = ns;
    }
  }
  int32_t query_min(int x0, int z0, int size) const {
    int level = (int)ctz32((uint32_t)size);
    int side = kSide >> level;
    return lv[level][(size_t)(z0 / size) * side + (x0 / size)];
  }
};

// ---------------------------------------------------------------------------
// Bottom-up sparse octree (svo/builder_svo.rs:69-534).
// ---------------------------------------------------------------------------
void build_svo(const uint8_t* material, const Prefix& prefix, const GroundMinMip& gmip,
               int32_t chunk_oy, int32_t dirt_depth, std::vector<vx_node>& out) {
  const int max_lvl = 6;
  std::vector<vx_node> packed_levels[7];

  // Leaves: one per solid voxel in linear order.
  std::vector<vx_node> child_nodes;
  std::vector<uint32_t> child_grid((size_t)kSide3, kInvalid);
  child_nodes.reserve(1 << 16);
  for (int i = 0; i < kSide3; ++i) {
    uint8_t m = material[i];
    if (m == VX_MAT_AIR) continue;
    int x = i & 63, z = (i >> 6) & 63, y = i >> 12;
    child_grid[i] = (uint32_t)child_nodes.size();
    child_nodes.push_back(make_leaf(x, y, z, 1, m));
  }

  int child_side = kSide, child_size = 1;
  std::vector<uint8_t> mask, collapse;
  std::vector<uint32_t> parent_grid;
  std::vector<vx_node> parent_nodes, packed;
  for (int step = 0; step < max_lvl; ++step) {
    const int parent_size = child_size * 2;
    const int parent_side = child_side / 2;
    const int pcells = parent_side * parent_side * parent_side;
    const int pbits = (int)ctz32((uint32_t)parent_side);
    const int pmask = parent_side - 1;
    mask.assign(pcells, 0);
    collapse.assign(pcells, 0);
    for (int pi = 0; pi < pcells; ++pi) {
      int px = pi & pmask, pz = (pi >> pbits) & pmask, py = pi >> (2 * pbits);
      uint8_t m = 0;
      for (int ci = 0; ci < 8; ++ci) {
        int cx = 2 * px + (ci & 1), cy = 2 * py + ((ci >> 1) & 1), cz = 2 * pz + ((ci >> 2) & 1);
        if (child_grid[lin_xyz(child_side, cx, cy, cz)] != kInvalid) m |= (uint8_t)(1u << ci);
      }
      mask[pi] = m;
      if (m == 0) continue;
      // Stone collapse: cube entirely below (min ground - dirt depth) and full.
      int ox = px * parent_size, oy = py * parent_size, oz = pz * parent_size;
      int32_t gmin = gmip.query_min(ox, oz, parent_size);
      int32_t y0 = chunk_oy + oy;
      int32_t y1 = y0 + parent_size - 1;
      if (y1 >= gmin - dirt_depth) continue;
      uint32_t full = (uint32_t)(parent_size * parent_size * parent_size);
      collapse[pi] = prefix.cube(ox, oy, oz, parent_size) == full ? 1 : 0;
    }
    parent_grid.assign(pcells, kInvalid);
    uint32_t pcount = 0;
    for (int pi = 0; pi < pcells; ++pi)
      if (mask[pi]) parent_grid[pi] = pcount++;
    parent_nodes.assign(pcount, make_leaf(0, 0, 0, parent_size, VX_MAT_AIR));
    packed.clear();
    for (int pi = 0; pi < pcells; ++pi) {
      uint8_t m = mask[pi];
      if (!m) continue;
      uint32_t pout = parent_grid[pi];
      int px = pi & pmask, pz = (pi >> pbits) & pmask, py = pi >> (2 * pbits);
      int ox = px * parent_size, oy = py * parent_size, oz = pz * parent_size;
      if (collapse[pi]) {
        parent_nodes[pout] = make_leaf(ox, oy, oz, parent_size, VX_MAT_STONE);
        continue;
      }
      int cx0 = 2 * px, cy0 = 2 * py, cz0 = 2 * pz;
      if (m == 0xFF) {
        const vx_node& c0 = child_nodes[child_grid[lin_xyz(child_side, cx0, cy0, cz0)]];
        if (c0.child_base == kInvalid) {
          bool ok = true;
          for (int ci = 1; ci < 8 && ok; ++ci) {
            uint32_t cidx = child_grid[lin_xyz(child_side, cx0 + (ci & 1), cy0 + ((ci >> 1) & 1),
                                               cz0 + ((ci >> 2) & 1))];
            if (cidx == kInvalid) { ok = false; break; }
            const vx_node& c = child_nodes[cidx];
            if (c.child_base != kInvalid || c.material != c0.material) ok = false;
          }
          if (ok) {
            parent_nodes[pout] = make_leaf(ox, oy, oz, parent_size, c0.material);
            continue;
          }
        }
      }
      uint32_t base = (uint32_t)packed.size();
      uint32_t cmask = 0;
      for (int ci = 0; ci < 8; ++ci) {
        if (!(m & (1u << ci))) continue;
        uint32_t cidx = child_grid[lin_xyz(child_side, cx0 + (ci & 1), cy0 + ((ci >> 1) & 1),
                                           cz0 + ((ci >> 2) & 1))];
        packed.push_back(child_nodes[cidx]);
        cmask |= 1u << ci;
      }
      vx_node pn;
      pn.child_base = base;
      pn.child_mask = cmask;
      pn.material = 0;
      pn.key = pack_key(kSide, ox, oy, oz, parent_size);
      parent_nodes[pout] = pn;
    }
    int child_level = (int)ctz32((uint32_t)(kSide / child_size));
    packed_levels[child_level].swap(packed);
    child_nodes.swap(parent_nodes);
    child_grid.swap(parent_grid);
    child_side = parent_side;
    child_size = parent_size;
  }

  vx_node root = child_nodes.empty() ? make_leaf(0, 0, 0, kSide, VX_MAT_AIR) : child_nodes[0];
  uint32_t bases[7] = {0, 0, 0, 0, 0, 0, 0};
  uint32_t cur = 1;
  for (int l = 1; l <= max_lvl; ++l) {
    bases[l] = cur;
    cur += (uint32_t)packed_levels[l].size();
  }
  if (root.child_base != kInvalid) root.child_base += bases[1];
  for (int l = 1; l < max_lvl; ++l)
    for (vx_node& n : packed_levels[l])
      if (n.child_base != kInvalid) n.child_base += bases[l + 1];
  out.clear();
  out.reserve(cur);
  out.push_back(root);
  for (int l = 1; l <= max_lvl; ++l) out.insert(out.end(), packed_levels[l].begin(), packed_levels[l].end());
}

// ---------------------------------------------------------------------------
// Ropes (svo/builder_ropes.rs:31-195): top-down, sibling first, else the
// parent's rope descended one level (or the leaf itself).
// ---------------------------------------------------------------------------
inline uint32_t child_of(const vx_node* nodes, uint32_t parent, uint32_t ci) {
  const vx_node& p = nodes[parent];
  uint32_t bit = 1u << ci;
  if (!(p.child_mask & bit)) return kInvalid;
  return p.child_base + (uint32_t)__builtin_popcount(p.child_mask & (bit - 1u));
}

inline uint32_t descend_one(const vx_node* nodes, uint32_t nei, uint32_t hx, uint32_t hy, uint32_t hz) {
  if (nei == kInvalid) return kInvalid;
  if (nodes[nei].child_base == kInvalid) return nei;
  return child_of(nodes, nei, hx | (hy << 1) | (hz << 2));
}

void ropes_rec(const vx_node* nodes, vx_node_ropes* ropes, uint32_t idx) {
  const vx_node& p = nodes[idx];
  if (p.child_base == kInvalid) return;
  const vx_node_ropes pr = ropes[idx];
  uint32_t kid[8];
  uint32_t rank = 0;
  for (uint32_t ci = 0; ci < 8; ++ci) kid[ci] = (p.child_mask & (1u << ci)) ? p.child_base + rank++ : kInvalid;
  for (uint32_t ci = 0; ci < 8; ++ci) {
    if (kid[ci] == kInvalid) continue;
    uint32_t hx = ci & 1u, hy = (ci >> 1) & 1u, hz = (ci >> 2) & 1u;
    vx_node_ropes r;
    // A face whose sibling lies inside the parent uses the sibling when present.
    r.px = (hx == 0 && kid[ci ^ 1u] != kInvalid) ? kid[ci ^ 1u] : descend_one(nodes, pr.px, 0, hy, hz);
    r.nx = (hx == 1 && kid[ci ^ 1u] != kInvalid) ? kid[ci ^ 1u] : descend_one(nodes, pr.nx, 1, hy, hz);
    r.py = (hy == 0 && kid[ci ^ 2u] != kInvalid) ? kid[ci ^ 2u] : descend_one(nodes, pr.py, hx, 0, hz);
    r.ny = (hy == 1 && kid[ci ^ 2u] != kInvalid) ? kid[ci ^ 2u] : descend_one(nodes, pr.ny, hx, 1, hz);
    r.pz = (hz == 0 && kid[ci ^ 4u] != kInvalid) ? kid[ci ^ 4u] : descend_one(nodes, pr.pz, hx, hy, 0);
    r.nz = (hz == 1 && kid[ci ^ 4u] != kInvalid) ? kid[ci ^ 4u] : descend_one(nodes, pr.nz, hx, hy, 1);
    r.pad0 = 0;
    r.pad1 = 0;
    ropes[kid[ci]] = r;
  }
  for (uint32_t ci = 0; ci < 8; ++ci)
    if (kid[ci] != kInvalid) ropes_rec(nodes, ropes, kid[ci]);
}

// ---------------------------------------------------------------------------
// Synthetic world.
// ---------------------------------------------------------------------------
using vxt::lattice_hash;
using vxt::mix32;

struct TreeInfo {
  bool present;
  int32_t trunk;
};

inline TreeInfo tree_at(const vx_world_params& p, int32_t wx, int32_t wz) {
  uint32_t h = lattice_hash(wx, wz, p.seed * 0x27d4eb2fu + 0x7EE5u);
  double u = (double)(h >> 8) * (1.0 / 16777216.0);
  TreeInfo t;
  t.present = u < (double)p.tree_density;
  uint32_t span = (uint32_t)(p.trunk_max - p.trunk_min + 1);
  t.trunk = p.trunk_min + (int32_t)(mix32(h ^ 0x5bd1e995u) % (span ? span : 1u));
  return t;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int vx_build_chunk(const uint8_t* material, const int32_t* ground, int32_t chunk_oy,
                   int32_t dirt_depth, vx_node* nodes, vx_node_ropes* ropes,
                   uint32_t node_cap, uint32_t* node_count,
                   uint32_t macro_words[16], uint32_t colinfo_words[2048]) {
  if (!material || !ground || !nodes || !ropes || !node_count || !macro_words || !colinfo_words)
    return vx_fail(VX_E_INVALID, "vx_build_chunk: null argument");
  try {
    // Column tops (svo/builder.rs:773-805) and colinfo words (742-771).
    uint8_t top_y[kSide2], top_m[kSide2];
    std::memset(top_y, 255, sizeof(top_y));
    std::memset(top_m, 0, sizeof(top_m));
    for (int ly = 0; ly < kSide; ++ly) {
      const uint8_t* slab = material + (size_t)ly * kSide2;
      for (int col = 0; col < kSide2; ++col)
        if (slab[col] != VX_MAT_AIR) {
          top_y[col] = (uint8_t)ly;
          top_m[col] = slab[col];
        }
    }
    std::memset(colinfo_words, 0, 2048 * sizeof(uint32_t));
    for (int idx = 0; idx < kSide2; ++idx) {
      uint32_t e = (uint32_t)top_y[idx] | ((uint32_t)top_m[idx] << 8);
      if (idx & 1) colinfo_words[idx >> 1] |= e << 16;
      else colinfo_words[idx >> 1] |= e;
    }
    // Summed-volume table and macro bits (svo/builder.rs:809-841).
    Prefix prefix;
    prefix.build(material);
    std::memset(macro_words, 0, 16 * sizeof(uint32_t));
    for (int mz = 0; mz < 8; ++mz)
      for (int my = 0; my < 8; ++my)
        for (int mx = 0; mx < 8; ++mx)
          if (prefix.cube(mx * 8, my * 8, mz * 8, 8) > 0) {
            uint32_t bit = (uint32_t)(mx + 8 * (my + 8 * mz));
            macro_words[bit >> 5] |= 1u << (bit & 31u);
          }
    GroundMinMip gmip;
    gmip.build(ground);
    std::vector<vx_node> out;
    build_svo(material, prefix, gmip, chunk_oy, dirt_depth, out);
    if (out.size() > node_cap) return vx_fail(VX_E_CAPACITY, "vx_build_chunk: node_cap too small");
    std::memcpy(nodes, out.data(), out.size() * sizeof(vx_node));
    for (size_t i = 0; i < out.size(); ++i) {
      vx_node_ropes r;
      r.px = r.nx = r.py = r.ny = r.pz = r.nz = kInvalid;
      r.pad0 = r.pad1 = 0;
      ropes[i] = r;
    }
    ropes_rec(nodes, ropes, 0);
    *node_count = (uint32_t)out.size();
  } catch (const std::bad_alloc&) {
    return vx_fail(VX_E_NOMEM, "vx_build_chunk: out of host memory");
  }
  return VX_OK;
}

void vx_world_default(vx_world_params* p) {
  if (!p) return;
  p->seed = 42;
  p->base_y = 96;
  p->amplitude = 48;
  p->dirt_depth = 4;
  p->tree_density = 0.01f;
  p->trunk_min = 5;
  p->trunk_max = 12;
  p->leaf_radius = 2;
  p->wavelength0 = 64;
}

int32_t vx_world_ground_height(const vx_world_params* p, int32_t wx, int32_t wz) {
  return vxt::ground_height(*p, wx, wz);
}

int vx_world_fill_chunk(const vx_world_params* p, int32_t cx, int32_t cy, int32_t cz,
                        uint8_t* material, int32_t* ground) {
  if (!p || !material || !ground) return vx_fail(VX_E_INVALID, "vx_world_fill_chunk: null argument");
  const int32_t ox = cx * kSide, oy = cy * kSide, oz = cz * kSide;
  const int R = std::max(0, p->leaf_radius);
  const int P = kSide + 2 * R;  // padded footprint
  std::vector<int32_t> g((size_t)P * P);
  for (int z = 0; z < P; ++z)
    for (int x = 0; x < P; ++x) g[(size_t)z * P + x] = vx_world_ground_height(p, ox - R + x, oz - R + z);
  for (int lz = 0; lz < kSide; ++lz)
    for (int lx = 0; lx < kSide; ++lx) ground[lz * kSide + lx] = g[(size_t)(lz + R) * P + (lx + R)];
  // Terrain: stone < ground - dirt_depth <= dirt < ground == grass.
  for (int ly = 0; ly < kSide; ++ly) {
    const int32_t wy = oy + ly;
    uint8_t* slab = material + (size_t)ly * kSide2;
    for (int col = 0; col < kSide2; ++col) {
      const int32_t gh = ground[col];
      uint8_t m;
      if (wy < gh) m = (wy >= gh - p->dirt_depth) ? VX_MAT_DIRT : VX_MAT_STONE;
      else if (wy == gh) m = VX_MAT_GRASS;
      else m = VX_MAT_AIR;
      slab[col] = (uint8_t)m;
    }
  }
  // Trees whose trunk or leaf blob reaches this chunk: trunks first (over air),
  // then leaf blobs (over remaining air), so the result is order independent.
  struct T { int32_t wx, wz, g, trunk; };
  std::vector<T> trees;
  for (int z = 0; z < P; ++z)
    for (int x = 0; x < P; ++x) {
      int32_t wx = ox - R + x, wz = oz - R + z;
      TreeInfo ti = tree_at(*p, wx, wz);
      if (ti.present) trees.push_back(T{wx, wz, g[(size_t)z * P + x], ti.trunk});
    }
  auto set_if_air = [&](int32_t wx, int32_t wy, int32_t wz, uint8_t m) {
    int lx = wx - ox, ly = wy - oy, lz = wz - oz;
    if (lx < 0 || ly < 0 || lz < 0 || lx >= kSide || ly >= kSide || lz >= kSide) return;
    uint8_t& dst = material[(size_t)ly * kSide2 + (size_t)lz * kSide + lx];
    if (dst == VX_MAT_AIR) dst = m;
  };
  for (const T& t : trees)
    for (int32_t y = t.g + 1; y <= t.g + t.trunk; ++y) set_if_air(t.wx, y, t.wz, VX_MAT_WOOD);
  for (const T& t : trees) {
    int32_t cy0 = t.g + t.trunk;
    for (int32_t dy = -R; dy <= R; ++dy)
      for (int32_t dz = -R; dz <= R; ++dz)
        for (int32_t dx = -R; dx <= R; ++dx) set_if_air(t.wx + dx, cy0 + dy, t.wz + dz, VX_MAT_LEAF);
  }
  return VX_OK;
}

// ---------------------------------------------------------------------------
// Host world
// ---------------------------------------------------------------------------
struct vx_host_world {
  vx_world_params params;
  int32_t origin[3];
  uint32_t dims[3];
  std::vector<vx_node> nodes;
  std::vector<vx_node_ropes> ropes;
  std::vector<vx_chunk_meta> meta;
  std::vector<uint32_t> macro;
  std::vector<uint32_t> colinfo;
  std::vector<uint32_t> grid;
  // Per-slot material copies (kept for edits).
  std::vector<std::vector<uint8_t>> mats;
  std::vector<std::vector<int32_t>> grounds;
};

namespace {
struct ChunkOut {
  std::vector<vx_node> nodes;
  std::vector<vx_node_ropes> ropes;
  uint32_t macro[16];
  uint32_t colinfo[2048];
  std::vector<uint8_t> mat;
  std::vector<int32_t> ground;
};

int build_one(const vx_world_params& p, int32_t cx, int32_t cy, int32_t cz, ChunkOut& out, bool fill) {
  if (fill) {
    out.mat.resize(kSide3);
    out.ground.resize(kSide2);
    int rc = vx_world_fill_chunk(&p, cx, cy, cz, out.mat.data(), out.ground.data());
    if (rc) return rc;
  }
  std::vector<vx_node> nodes(VX_MAX_CHUNK_NODES);
  std::vector<vx_node_ropes> ropes(VX_MAX_CHUNK_NODES);
  uint32_t n = 0;
  int rc = vx_build_chunk(out.mat.data(), out.ground.data(), cy * kSide, p.dirt_depth, nodes.data(),
                          ropes.data(), VX_MAX_CHUNK_NODES, &n, out.macro, out.colinfo);
  if (rc) return rc;
  out.nodes.assign(nodes.begin(), nodes.begin() + n);
  out.ropes.assign(ropes.begin(), ropes.begin() + n);
  return VX_OK;
}

void pack_world(vx_host_world* w, std::vector<ChunkOut>& outs) {
  const uint32_t nchunks = (uint32_t)outs.size();
  size_t total = 0;
  for (auto& o : outs) total += o.nodes.size();
  w->nodes.resize(total);
  w->ropes.resize(total);
  w->meta.resize(nchunks);
  w->macro.resize((size_t)nchunks * 16);
  w->colinfo.resize((size_t)nchunks * 2048);
  w->grid.assign((size_t)w->dims[0] * w->dims[1] * w->dims[2], kInvalid);
  w->mats.resize(nchunks);
  w->grounds.resize(nchunks);
  size_t base = 0;
  for (uint32_t slot = 0; slot < nchunks; ++slot) {
    ChunkOut& o = outs[slot];
    const uint32_t ix = slot % w->dims[0];
    const uint32_t iy = (slot / w->dims[0]) % w->dims[1];
    const uint32_t iz = slot / (w->dims[0] * w->dims[1]);
    std::copy(o.nodes.begin(), o.nodes.end(), w->nodes.begin() + base);
    std::copy(o.ropes.begin(), o.ropes.end(), w->ropes.begin() + base);
    std::memcpy(&w->macro[(size_t)slot * 16], o.macro, sizeof(o.macro));
    std::memcpy(&w->colinfo[(size_t)slot * 2048], o.colinfo, sizeof(o.colinfo));
    vx_chunk_meta m;
    std::memset(&m, 0, sizeof(m));
    m.origin[0] = (w->origin[0] + (int32_t)ix) * kSide;
    m.origin[1] = (w->origin[1] + (int32_t)iy) * kSide;
    m.origin[2] = (w->origin[2] + (int32_t)iz) * kSide;
    m.node_base = (uint32_t)base;
    m.node_count = (uint32_t)o.nodes.size();
    m.macro_base = slot * 16;
    m.colinfo_base = slot * 2048;
    bool empty = true;
    for (int i = 0; i < 16; ++i) empty = empty && o.macro[i] == 0;
    m.macro_empty = empty ? 1u : 0u;
    w->meta[slot] = m;
    w->grid[slot] = slot;  // grid index == slot for a fully resident box
    w->mats[slot].swap(o.mat);
    w->grounds[slot].swap(o.ground);
    base += o.nodes.size();
  }
}
}  // namespace

int vx_host_world_build(const vx_world_params* p, const int32_t grid_origin[3],
                        const uint32_t grid_dims[3], uint32_t threads, vx_host_world** out) {
  if (!p || !grid_origin || !grid_dims || !out) return vx_fail(VX_E_INVALID, "vx_host_world_build: null argument");
  uint64_t n64 = (uint64_t)grid_dims[0] * grid_dims[1] * grid_dims[2];
  if (n64 == 0 || n64 > (1u << 20)) return vx_fail(VX_E_INVALID, "vx_host_world_build: bad grid dims");
  vx_host_world* w = new (std::nothrow) vx_host_world();
  if (!w) return vx_fail(VX_E_NOMEM, "vx_host_world_build: out of memory");
  w->params = *p;
  for (int i = 0; i < 3; ++i) {
    w->origin[i] = grid_origin[i];
    w->dims[i] = grid_dims[i];
  }
  const uint32_t n = (uint32_t)n64;
  std::vector<ChunkOut> outs(n);
  if (threads == 0) threads = std::max(1u, std::thread::hardware_concurrency());
  threads = std::min(threads, n);
  std::atomic<uint32_t> next{0};
  std::atomic<int> err{0};
  auto worker = [&]() {
    for (;;) {
      uint32_t i = next.fetch_add(1);
      if (i >= n || err.load()) return;
      int32_t cx = grid_origin[0] + (int32_t)(i % grid_dims[0]);
      int32_t cy = grid_origin[1] + (int32_t)((i / grid_dims[0]) % grid_dims[1]);
      int32_t cz = grid_origin[2] + (int32_t)(i / (grid_dims[0] * grid_dims[1]));
      int rc = build_one(*p, cx, cy, cz, outs[i], true);
      if (rc) err.store(rc);
    }
  };
  std::vector<std::thread> pool;
  for (uint32_t t = 1; t < threads; ++t) pool.emplace_back(worker);
  worker();
  for (auto& th : pool) th.join();
  if (err.load()) {
    delete w;
    return err.load();
  }
  pack_world(w, outs);
  *out = w;
  return VX_OK;
}

void vx_host_world_free(vx_host_world* w) { delete w; }

int vx_host_world_get(const vx_host_world* w, vx_host_world_view* v) {
  if (!w || !v) return vx_fail(VX_E_INVALID, "vx_host_world_get: null argument");
  v->nodes = w->nodes.data();
  v->node_count = w->nodes.size();
  v->ropes = w->ropes.data();
  v->meta = w->meta.data();
  v->chunk_count = (uint32_t)w->meta.size();
  v->macro_words = w->macro.data();
  v->colinfo_words = w->colinfo.data();
  v->grid = w->grid.data();
  v->grid_cells = (uint32_t)w->grid.size();
  for (int i = 0; i < 3; ++i) {
    v->grid_origin[i] = w->origin[i];
    v->grid_dims[i] = w->dims[i];
  }
  return VX_OK;
}

int vx_host_world_edit(vx_host_world* w, const int32_t* xyz, const uint8_t* mats, uint32_t n,
                       uint32_t* rebuilt_slots, uint32_t* n_rebuilt) {
  if (!w || (!xyz && n) || (!mats && n) || !n_rebuilt) return vx_fail(VX_E_INVALID, "vx_host_world_edit: null argument");
  std::vector<uint32_t> touched;
  for (uint32_t i = 0; i < n; ++i) {
    int32_t wx = xyz[3 * i], wy = xyz[3 * i + 1], wz = xyz[3 * i + 2];
    int32_t cx = (int32_t)std::floor((double)wx / kSide), cy = (int32_t)std::floor((double)wy / kSide),
            cz = (int32_t)std::floor((double)wz / kSide);
    int32_t ix = cx - w->origin[0], iy = cy - w->origin[1], iz = cz - w->origin[2];
    if (ix < 0 || iy < 0 || iz < 0 || ix >= (int32_t)w->dims[0] || iy >= (int32_t)w->dims[1] ||
        iz >= (int32_t)w->dims[2])
      continue;
    uint32_t slot = w->grid[((size_t)iz * w->dims[1] + iy) * w->dims[0] + ix];
    if (slot == kInvalid) continue;
    int lx = wx - cx * kSide, ly = wy - cy * kSide, lz = wz - cz * kSide;
    // Edit index uses the material layout (world/edits.rs:40-45).
    w->mats[slot][(size_t)ly * kSide2 + (size_t)lz * kSide + lx] = mats[i];
    if (std::find(touched.begin(), touched.end(), slot) == touched.end()) touched.push_back(slot);
  }
  // Rebuild touched chunks; node ranges change size, so the whole arena is
  // repacked (host side only; the device upload is the caller's region write).
  std::vector<ChunkOut> outs(w->meta.size());
  for (uint32_t slot = 0; slot < w->meta.size(); ++slot) {
    ChunkOut& o = outs[slot];
    o.mat.swap(w->mats[slot]);
    o.ground.swap(w->grounds[slot]);
    bool t = std::find(touched.begin(), touched.end(), slot) != touched.end();
    if (t) {
      int32_t cy = w->meta[slot].origin[1] / kSide;
      int rc = build_one(w->params, 0, cy, 0, o, false);
      if (rc) return rc;
    } else {
      const vx_chunk_meta& m = w->meta[slot];
      o.nodes.assign(w->nodes.begin() + m.node_base, w->nodes.begin() + m.node_base + m.node_count);
      o.ropes.assign(w->ropes.begin() + m.node_base, w->ropes.begin() + m.node_base + m.node_count);
      std::memcpy(o.macro, &w->macro[(size_t)slot * 16], sizeof(o.macro));
      std::memcpy(o.colinfo, &w->colinfo[(size_t)slot * 2048], sizeof(o.colinfo));
    }
  }
  pack_world(w, outs);
  *n_rebuilt = (uint32_t)touched.size();
  if (rebuilt_slots)
    for (size_t i = 0; i < touched.size(); ++i) rebuilt_slots[i] = touched[i];
  return VX_OK;
}

// ---------------------------------------------------------------------------
// Camera (app/app.rs:411-511, app/camera.rs:95-111, glam 0.27 conventions).
// ---------------------------------------------------------------------------
namespace {
struct M4 {
  double m[16];  // column-major: m[c*4 + r]
  double& at(int r, int c) { return m[c * 4 + r]; }
  double at(int r, int c) const { return m[c * 4 + r]; }
};
M4 mul(const M4& a, const M4& b) {
  M4 r;
  for (int c = 0; c < 4; ++c)
    for (int i = 0; i < 4; ++i) {
      double s = 0;
      for (int k = 0; k < 4; ++k) s += a.at(i, k) * b.at(k, c);
      r.at(i, c) = s;
    }
  return r;
}
M4 inverse(const M4& a) {
  // Gauss-Jordan with partial pivoting.
  double w[4][8];
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 8; ++c) w[r][c] = c < 4 ? a.at(r, c) : (c - 4 == r ? 1.0 : 0.0);
  for (int c = 0; c < 4; ++c) {
    int piv = c;
    for (int r = c + 1; r < 4; ++r)
      if (std::fabs(w[r][c]) > std::fabs(w[piv][c])) piv = r;
    for (int k = 0; k < 8; ++k) std::swap(w[c][k], w[piv][k]);
    double d = w[c][c];
    for (int k = 0; k < 8; ++k) w[c][k] /= d;
    for (int r = 0; r < 4; ++r) {
      if (r == c) continue;
      double f = w[r][c];
      for (int k = 0; k < 8; ++k) w[r][k] -= f * w[c][k];
    }
  }
  M4 out;
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) out.at(r, c) = w[r][c + 4];
  return out;
}
double halton(uint32_t index, uint32_t base) {
  double f = 1.0, r = 0.0;
  while (index > 0) {
    f /= (double)base;
    r += f * (double)(index % base);
    index /= base;
  }
  return r;
}
void store(float* dst, const M4& m) {
  for (int i = 0; i < 16; ++i) dst[i] = (float)m.m[i];
}
}  // namespace

int vx_camera_build(const vx_pose* pose, uint32_t render_w, uint32_t render_h, uint32_t present_w,
                    uint32_t present_h, uint32_t frame_index, float time_s, float wind_strength,
                    float fog_density, const int32_t grid_origin[3], const uint32_t grid_dims[3],
                    uint32_t chunk_count, const float* prev_view_proj, uint32_t flags, vx_camera* out) {
  if (!pose || !out || !grid_origin || !grid_dims || !render_w || !render_h || !present_w || !present_h)
    return vx_fail(VX_E_INVALID, "vx_camera_build: bad argument");
  const double yaw = pose->yaw, pitch = pose->pitch;
  double f[3] = {std::sin(yaw) * std::cos(pitch), std::sin(pitch), std::cos(yaw) * std::cos(pitch)};
  double fl = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
  for (double& v : f) v /= fl;
  // look_at_rh(eye, eye + f, +Y)
  double up[3] = {0, 1, 0};
  double s[3] = {f[1] * up[2] - f[2] * up[1], f[2] * up[0] - f[0] * up[2], f[0] * up[1] - f[1] * up[0]};
  double sl = std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
  for (double& v : s) v /= sl;
  double u[3] = {s[1] * f[2] - s[2] * f[1], s[2] * f[0] - s[0] * f[2], s[0] * f[1] - s[1] * f[0]};
  const double e[3] = {pose->pos[0], pose->pos[1], pose->pos[2]};
  M4 view;
  std::memset(view.m, 0, sizeof(view.m));
  for (int i = 0; i < 3; ++i) {
    view.at(0, i) = s[i];
    view.at(1, i) = u[i];
    view.at(2, i) = -f[i];
  }
  view.at(0, 3) = -(s[0] * e[0] + s[1] * e[1] + s[2] * e[2]);
  view.at(1, 3) = -(u[0] * e[0] + u[1] * e[1] + u[2] * e[2]);
  view.at(2, 3) = (f[0] * e[0] + f[1] * e[1] + f[2] * e[2]);
  view.at(3, 3) = 1.0;
  // perspective_rh(fovy, aspect, near, far); aspect of the present surface.
  const double aspect = (double)present_w / (double)present_h;
  const double h = 1.0 / std::tan(0.5 * pose->fovy), w = h / aspect;
  const double zn = pose->z_near, zf = pose->z_far, r = zf / (zn - zf);
  M4 proj;
  std::memset(proj.m, 0, sizeof(proj.m));
  proj.at(0, 0) = w;
  proj.at(1, 1) = h;
  proj.at(2, 2) = r;
  proj.at(3, 2) = -1.0;
  proj.at(2, 3) = r * zn;
  M4 view_inv = inverse(view), proj_inv = inverse(proj);
  M4 vp = mul(proj, view);

  double jx = 0.0, jy = 0.0;
  if (flags & VX_CAM_JITTER) {
    jx = halton(frame_index + 1u, 2) - 0.5;
    jy = halton(frame_index + 1u, 3) - 0.5;
  }
  auto ray_dir_at = [&](double px, double py, double outv[3]) {
    double ndc[4] = {2.0 * px / render_w - 1.0, 1.0 - 2.0 * py / render_h, 1.0, 1.0};
    double vpos[4];
    for (int rr = 0; rr < 4; ++rr) {
      vpos[rr] = 0;
      for (int k = 0; k < 4; ++k) vpos[rr] += proj_inv.at(rr, k) * ndc[k];
    }
    double vd[4] = {vpos[0] / vpos[3], vpos[1] / vpos[3], vpos[2] / vpos[3], 0.0};
    for (int rr = 0; rr < 3; ++rr) {
      outv[rr] = 0;
      for (int k = 0; k < 4; ++k) outv[rr] += view_inv.at(rr, k) * vd[k];
    }
  };
  double r00[3], r10[3], r01[3];
  ray_dir_at(0.5 + jx, 0.5 + jy, r00);
  ray_dir_at(1.5 + jx, 0.5 + jy, r10);
  ray_dir_at(0.5 + jx, 1.5 + jy, r01);

  std::memset(out, 0, sizeof(*out));
  store(out->view_inv, view_inv);
  store(out->proj_inv, proj_inv);
  store(out->view_proj, vp);
  if (prev_view_proj) std::memcpy(out->prev_view_proj, prev_view_proj, 16 * sizeof(float));
  else store(out->prev_view_proj, vp);
  out->cam_pos[0] = pose->pos[0];
  out->cam_pos[1] = pose->pos[1];
  out->cam_pos[2] = pose->pos[2];
  out->cam_pos[3] = 1.0f;
  for (int i = 0; i < 3; ++i) {
    out->ray00[i] = (float)r00[i];
    out->ray_dx[i] = (float)(r10[i] - r00[i]);
    out->ray_dy[i] = (float)(r01[i] - r00[i]);
  }
  out->chunk_size = VX_CHUNK_SIZE;
  out->chunk_count = chunk_count;
  out->max_steps = 128;  // clamp(CHUNK_SIZE * 2, 64, 256)
  out->frame_index = frame_index;
  out->voxel_params[0] = 0.1f;
  out->voxel_params[1] = time_s;
  out->voxel_params[2] = wind_strength;
  out->voxel_params[3] = fog_density;
  for (int i = 0; i < 3; ++i) {
    out->grid_origin_chunk[i] = grid_origin[i];
    out->grid_dims[i] = grid_dims[i];
  }
  out->render_present_px[0] = render_w;
  out->render_present_px[1] = render_h;
  out->render_present_px[2] = present_w;
  out->render_present_px[3] = present_h;
  return VX_OK;
}

// ---------------------------------------------------------------------------
// Clipmap (clipmap.rs:62-122): snapped origin = floor((cam - res/2*cell)/cell),
// heights at texel centres, storage = (logical + origin_cell) mod res.
// ---------------------------------------------------------------------------
int vx_clipmap_build(const vx_world_params* p, float cam_x, float cam_z, uint32_t levels, uint32_t res,
                     float base_cell_m, vx_clipmap* out, float* heights) {
  if (!p || !out || !heights || levels > VX_CLIP_LEVELS_MAX || res == 0 || res > 4096 || !(base_cell_m > 0.0f) ||
      !std::isfinite(base_cell_m))
    return vx_fail(VX_E_INVALID, "vx_clipmap_build: bad argument");
  std::memset(out, 0, sizeof(*out));
  out->levels = levels;
  out->res = res;
  out->base_cell_m = base_cell_m;
  for (uint32_t l = 0; l < VX_CLIP_LEVELS_MAX; ++l) {
    const float cell = base_cell_m * (float)(1u << l);
    const float half = (float)res * 0.5f;
    const int32_t ocx = vxt::sat_i32(std::floor((cam_x - half * cell) / cell));
    const int32_t ocz = vxt::sat_i32(std::floor((cam_z - half * cell) / cell));
    const float oxm = (float)ocx * cell, ozm = (float)ocz * cell;
    int32_t offx = ocx % (int32_t)res, offz = ocz % (int32_t)res;
    if (offx < 0) offx += (int32_t)res;
    if (offz < 0) offz += (int32_t)res;
    out->level[l][0] = oxm;
    out->level[l][1] = ozm;
    out->level[l][2] = cell;
    out->level[l][3] = 0.0f;
    out->offset[l][0] = (uint32_t)offx;
    out->offset[l][1] = (uint32_t)offz;
    if (l >= levels) continue;
    float* layer = heights + (size_t)l * res * res;
    for (uint32_t tz = 0; tz < res; ++tz) {
      const uint32_t sz = (tz + (uint32_t)offz) % res;
      for (uint32_t tx = 0; tx < res; ++tx)
        layer[(size_t)sz * res + (tx + (uint32_t)offx) % res] = vxt::clip_texel_height(*p, oxm, ozm, cell, tx, tz);
    }
  }
  return VX_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Incremental clipmap (clipmap.rs:44-60 state, 182-577 update, 580-585
// invalidate).  Per level the state holds the committed origin cell and the
// toroidal offset (storage = (logical + offset) mod res); an update scrolls the
// offset by the origin's cell delta and plans the texels that entered the
// level as storage rectangles, in the reference's upload order: a full level on
// first touch or a jump of res cells or more, else the new columns (split in
// two where they wrap) then the new rows.
// ---------------------------------------------------------------------------
struct vx_clipmap_state {
  uint32_t levels = 0, res = 0, flags = 0;
  float base_cell = 1.0f, min_interval = 0.0f;
  int32_t cell[VX_CLIP_LEVELS_MAX][2];
  uint32_t off[VX_CLIP_LEVELS_MAX][2];
  float last_time = -INFINITY;

  void invalidate() {
    for (uint32_t l = 0; l < VX_CLIP_LEVELS_MAX; ++l) {
      cell[l][0] = cell[l][1] = INT32_MIN;
      off[l][0] = off[l][1] = 0;
    }
    last_time = -INFINITY;
  }
};

namespace {

// Logical texel of the camera in a level whose origin cell is `o`.
inline int32_t cam_texel(float cam, int32_t o, float cell) { return vxt::sat_i32(std::floor((cam - (float)o * cell) / cell)); }

inline uint32_t wrap_mod(int64_t v, uint32_t m) {
  int64_t r = v % (int64_t)m;
  return (uint32_t)(r < 0 ? r + (int64_t)m : r);
}

struct LevelPlan {
  float cell, ox, oz;  // new origin (metres) and cell size
  int32_t cx, cz;      // new origin cell
  int64_t dx, dz;      // cell delta from the committed origin (0 on first touch)
  bool first, moved;
};

}  // namespace

extern "C" {

int vx_clipmap_state_create(uint32_t levels, uint32_t res, float base_cell_m, float min_update_interval_s,
                            uint32_t flags, vx_clipmap_state** out) {
  if (!out || levels == 0 || levels > VX_CLIP_LEVELS_MAX || res < 2 * VX_CLIP_EDGE_GUARD + 1 || res > 4096 ||
      !(base_cell_m > 0.0f) || !std::isfinite(base_cell_m) || (flags & ~VX_CLIP_WRAP_FIX))
    return vx_fail(VX_E_INVALID, "vx_clipmap_state_create: bad argument");
  vx_clipmap_state* s = new (std::nothrow) vx_clipmap_state();
  if (!s) return vx_fail(VX_E_NOMEM, "vx_clipmap_state_create: host allocation");
  s->levels = levels;
  s->res = res;
  s->base_cell = base_cell_m;
  s->min_interval = min_update_interval_s;
  s->flags = flags;
  s->invalidate();
  *out = s;
  return VX_OK;
}

void vx_clipmap_state_destroy(vx_clipmap_state* s) { delete s; }

int vx_clipmap_invalidate(vx_clipmap_state* s) {
  if (!s) return vx_fail(VX_E_INVALID, "vx_clipmap_invalidate: null state");
  s->invalidate();
  return VX_OK;
}

int vx_clipmap_update(vx_clipmap_state* s, float cam_x, float cam_z, float time_s, vx_clipmap* params,
                      vx_clip_rect* rects, uint32_t cap, uint32_t* n_rects) {
  if (!s || !params || !n_rects || !rects) return vx_fail(VX_E_INVALID, "vx_clipmap_update: null argument");
  if (cap < 4 * s->levels)
    return vx_fail(VX_E_INVALID, "vx_clipmap_update: room for %u rectangles, an update may emit %u", cap, 4 * s->levels);
  const uint32_t res = s->res;
  const int32_t guard = VX_CLIP_EDGE_GUARD, hi = (int32_t)res - VX_CLIP_EDGE_GUARD;
  const float half = (float)res * 0.5f;

  // Where every level wants to be, and whether any level is new, moved or
  // has the camera within the guard band of its committed window.
  LevelPlan plan[VX_CLIP_LEVELS_MAX];
  bool any_first = false, any_moved = false, any_force = false;
  for (uint32_t l = 0; l < s->levels; ++l) {
    LevelPlan& q = plan[l];
    q.cell = s->base_cell * (float)(1u << l);
    q.cx = vxt::sat_i32(std::floor((cam_x - half * q.cell) / q.cell));
    q.cz = vxt::sat_i32(std::floor((cam_z - half * q.cell) / q.cell));
    q.ox = (float)q.cx * q.cell;
    q.oz = (float)q.cz * q.cell;
    q.first = s->cell[l][0] == INT32_MIN;
    q.dx = q.first ? 0 : (int64_t)q.cx - s->cell[l][0];
    q.dz = q.first ? 0 : (int64_t)q.cz - s->cell[l][1];
    q.moved = q.first || q.dx != 0 || q.dz != 0;
    bool force = q.first;
    if (!q.first) {
      const int32_t tx = cam_texel(cam_x, s->cell[l][0], q.cell), tz = cam_texel(cam_z, s->cell[l][1], q.cell);
      force = tx < guard || tz < guard || tx >= hi || tz >= hi;
    }
    any_first |= q.first;
    any_moved |= q.moved;
    any_force |= force;
  }
  // One cadence gate for all levels.
  const bool cadence_ok = any_first || (time_s - s->last_time) >= s->min_interval;
  const bool update = any_first || any_force || (cadence_ok && any_moved);
  if (update) s->last_time = time_s;

  uint32_t n = 0;
  auto emit = [&](uint32_t l, uint32_t x, uint32_t y, uint32_t w, uint32_t h, uint32_t lx0, uint32_t lz0) {
    vx_clip_rect& r = rects[n++];
    r.level = l;
    r.x = x;
    r.y = y;
    r.w = w;
    r.h = h;
    r.lx0 = lx0;
    r.lz0 = lz0;
    r.res = res;
    r.origin_x_m = plan[l].ox;
    r.origin_z_m = plan[l].oz;
    r.cell_m = plan[l].cell;
    r.pad = 0;
  };
  // A strip of `k` new logical lines starting at logical line `first` along
  // one axis, at storage line (first + off) mod res, split where it wraps.
  // `along` is the logical start of the strip's full-length axis at storage 0.
  auto strip = [&](uint32_t l, bool cols, uint32_t first, uint32_t k, uint32_t off, uint32_t along) {
    const uint32_t s0 = (first + off) % res;
    const uint32_t a = s0 + k <= res ? k : res - s0;
    if (cols) {
      emit(l, s0, 0, a, res, first, along);
      if (a < k) emit(l, 0, 0, k - a, res, first + a, along);
    } else {
      emit(l, 0, s0, res, a, along, first);
      if (a < k) emit(l, 0, 0, res, k - a, along, first + a);
    }
  };

  const bool fix = (s->flags & VX_CLIP_WRAP_FIX) != 0;
  for (uint32_t l = 0; l < s->levels; ++l) {
    const LevelPlan& q = plan[l];
    if (update && q.moved) {
      const uint64_t adx = (uint64_t)(q.dx < 0 ? -q.dx : q.dx), adz = (uint64_t)(q.dz < 0 ? -q.dz : q.dz);
      if (q.first || adx >= res || adz >= res) {
        s->off[l][0] = s->off[l][1] = 0;
        emit(l, 0, 0, res, res, 0, 0);
      } else {
        s->off[l][0] = wrap_mod((int64_t)s->off[l][0] + q.dx, res);
        s->off[l][1] = wrap_mod((int64_t)s->off[l][1] + q.dz, res);
        // A strip spans the whole other axis.  The reference writes that axis
        // from storage 0 in logical order (clipmap.rs:376-377,470-471), which
        // agrees with the sampler's storage = (logical + offset) mod res only
        // while that axis' offset is 0; VX_CLIP_WRAP_FIX rotates it by the offset.
        const uint32_t along_z = fix ? (res - s->off[l][1]) % res : 0;
        const uint32_t along_x = fix ? (res - s->off[l][0]) % res : 0;
        if (q.dx != 0) strip(l, true, q.dx > 0 ? res - (uint32_t)adx : 0, (uint32_t)adx, s->off[l][0], along_z);
        if (q.dz != 0) strip(l, false, q.dz > 0 ? res - (uint32_t)adz : 0, (uint32_t)adz, s->off[l][1], along_x);
      }
      s->cell[l][0] = q.cx;
      s->cell[l][1] = q.cz;
    }
  }

  // The committed state: heights and uniform stay consistent (clipmap.rs:560-575).
  std::memset(params, 0, sizeof(*params));
  params->levels = s->levels;
  params->res = res;
  params->base_cell_m = s->base_cell;
  for (uint32_t l = 0; l < VX_CLIP_LEVELS_MAX; ++l) {
    if (l >= s->levels) {
      params->level[l][2] = 1.0f;
      continue;
    }
    const LevelPlan& q = plan[l];
    const bool committed = s->cell[l][0] != INT32_MIN;
    params->level[l][0] = committed ? (float)s->cell[l][0] * q.cell : q.ox;
    params->level[l][1] = committed ? (float)s->cell[l][1] * q.cell : q.oz;
    params->level[l][2] = q.cell;
    params->offset[l][0] = s->off[l][0];
    params->offset[l][1] = s->off[l][1];
  }
  *n_rects = n;
  return VX_OK;
}

int vx_clipmap_fill_rects(const vx_world_params* p, const vx_clip_rect* rects, uint32_t n, float* data) {
  if (!p || (n && (!rects || !data))) return vx_fail(VX_E_INVALID, "vx_clipmap_fill_rects: null argument");
  for (uint32_t i = 0; i < n; ++i) {
    const vx_clip_rect& r = rects[i];
    if (!r.res || r.w > r.res || r.h > r.res || r.lx0 >= r.res || r.lz0 >= r.res)
      return vx_fail(VX_E_INVALID, "vx_clipmap_fill_rects: rectangle %u out of range", i);
  }
  for (uint32_t i = 0; i < n; ++i) {
    const vx_clip_rect& r = rects[i];
    for (uint32_t j = 0; j < r.h; ++j) {
      const uint32_t lz = (r.lz0 + j) % r.res;
      for (uint32_t k = 0; k < r.w; ++k)
        *data++ = vxt::clip_texel_height(*p, r.origin_x_m, r.origin_z_m, r.cell_m, (r.lx0 + k) % r.res, lz);
    }
  }
  return VX_OK;
}

}  // extern "C"
