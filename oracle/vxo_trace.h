// vxo_trace.h — oracle restatement of ray/chunk_trace.wgsl (tile candidates,
// scene/chunk traversal, macro DDA, rope leaf traversal, grass column probe),
// ray/shadows.wgsl and ray/light.wgsl.
//
// TEST INFRASTRUCTURE ONLY (the checker of the HIP passes; never shipped).
#pragma once
#include "vxo_core.h"

namespace vxo {

// HitGeom, chunk_trace.wgsl:15-32, plus the hit identity (slot, leaf node,
// voxel) that the reference does not carry.
struct HitGeom {
  uint32_t hit, mat;
  float t;
  v3 n;
  v3 root_bmin;
  float root_size;
  uint32_t node_base, macro_base;
  uint32_t slot, node;
  int32_t vox[3];
};
static inline HitGeom miss_hitgeom() {
  HitGeom h;
  h.hit = 0;
  h.mat = MAT_AIR;
  h.t = BIG_F32;
  h.n = V3s(0.0f);
  h.root_bmin = V3s(0.0f);
  h.root_size = 0.0f;
  h.node_base = 0;
  h.macro_base = INVALID_U32;
  h.slot = INVALID_U32;
  h.node = INVALID_U32;
  h.vox[0] = h.vox[1] = h.vox[2] = 0;
  return h;
}
struct ChunkTraceResult { HitGeom hit; bool anchor_valid; uint32_t anchor_key; };

// point_in_cube, chunk_trace.wgsl:54-59
static inline bool point_in_cube(v3 p, v3 bmin, float size) {
  v3 bmax = bmin + V3s(size);
  return (p.x >= bmin.x && p.x < bmax.x) && (p.y >= bmin.y && p.y < bmax.y) && (p.z >= bmin.z && p.z < bmax.z);
}
// node_level/x/y/z, chunk_trace.wgsl:118-121
static inline uint32_t node_level(uint32_t k) { return (k >> 18u) & 7u; }
static inline uint32_t node_x(uint32_t k) { return k & 63u; }
static inline uint32_t node_y(uint32_t k) { return (k >> 6u) & 63u; }
static inline uint32_t node_z(uint32_t k) { return (k >> 12u) & 63u; }
// node_cube_from_key, chunk_trace.wgsl:123-128
static inline v4 node_cube_from_key(v3 root_bmin, float root_size, uint32_t key) {
  uint32_t lvl = node_level(key);
  float s = root_size / (float)(1u << lvl);
  v3 bmin = root_bmin + V3((float)node_x(key), (float)node_y(key), (float)node_z(key)) * s;
  return v4{bmin.x, bmin.y, bmin.z, s};
}
struct AnchorNode { bool valid; uint32_t idx; v3 bmin; float size; uint32_t key; };
static inline AnchorNode anchor_none() { return AnchorNode{false, 0u, V3s(0.0f), 0.0f, INVALID_U32}; }
// anchor_from_key, chunk_trace.wgsl:138-191
static inline AnchorNode anchor_from_key(const Ctx& C, uint32_t node_base, v3 root_bmin, float root_size, uint32_t key) {
  if (key == INVALID_U32) return anchor_none();
  uint32_t lvl = node_level(key), md = chunk_max_depth(C);
  lvl = lvl < md ? lvl : md;
  uint32_t tx = node_x(key), ty = node_y(key), tz = node_z(key);
  uint32_t idx = 0;
  v3 bmin = root_bmin;
  float size = root_size;
  for (uint32_t d = 0; d < 32u; ++d) {
    if (d >= lvl) {
      uint32_t nk = ld_node(C, (uint64_t)node_base + idx).key;
      return AnchorNode{true, idx, bmin, size, nk};
    }
    vx_node n = ld_node(C, (uint64_t)node_base + idx);
    if (n.child_base == LEAF_U32) return AnchorNode{true, idx, bmin, size, n.key};
    uint32_t shift = (lvl - 1u) - d;
    uint32_t hx = (tx >> shift) & 1u, hy = (ty >> shift) & 1u, hz = (tz >> shift) & 1u;
    uint32_t ci = hx | (hy << 1u) | (hz << 2u);
    uint32_t bit = 1u << ci;
    if ((n.child_mask & bit) == 0u) return AnchorNode{true, idx, bmin, size, n.key};
    float half = size * 0.5f;
    bmin = bmin + V3(hx ? half : 0.0f, hy ? half : 0.0f, hz ? half : 0.0f);
    size = half;
    idx = n.child_base + child_rank(n.child_mask, ci);
  }
  return AnchorNode{true, idx, bmin, size, ld_node(C, (uint64_t)node_base + idx).key};
}
// LeafState, chunk_trace.wgsl:281-293
struct LeafState {
  bool has_node;
  uint32_t idx_local;
  v3 bmin;
  float size;
  uint32_t mat;
  bool has_anchor;
  uint32_t anchor_idx;
  v3 anchor_bmin;
  float anchor_size;
};
// anchor_from_leaf, chunk_trace.wgsl:193-203
static inline AnchorNode anchor_from_leaf(const Ctx& C, const LeafState& leaf, uint32_t node_base) {
  if (leaf.has_node) {
    uint32_t k = ld_node(C, (uint64_t)node_base + leaf.idx_local).key;
    return AnchorNode{true, leaf.idx_local, leaf.bmin, leaf.size, k};
  }
  if (leaf.has_anchor) {
    uint32_t k = ld_node(C, (uint64_t)node_base + leaf.anchor_idx).key;
    return AnchorNode{true, leaf.anchor_idx, leaf.anchor_bmin, leaf.anchor_size, k};
  }
  return anchor_none();
}
// rope_next_local, chunk_trace.wgsl:217-225
static inline uint32_t rope_next_local(const Ctx& C, uint32_t node_base, uint32_t local_idx, uint32_t face) {
  vx_node_ropes r = ld_ropes(C, (uint64_t)node_base + local_idx);
  if (face == 0u) return r.px;
  if (face == 1u) return r.nx;
  if (face == 2u) return r.py;
  if (face == 3u) return r.ny;
  if (face == 4u) return r.pz;
  return r.nz;
}
struct ExitFaceRes { uint32_t face, amb; };
// exit_face_from_slab_safe, chunk_trace.wgsl:233-272
static inline ExitFaceRes exit_face_from_slab_safe(v3 rd, const CubeSlab& slab) {
  float te = slab.t_exit;
  float eps = 1e-5f * mx(1.0f, std::fabs(te));
  bool bx = (std::fabs(rd.x) >= EPS_INV) && (std::fabs(slab.tmaxv.x - te) <= eps);
  bool by = (std::fabs(rd.y) >= EPS_INV) && (std::fabs(slab.tmaxv.y - te) <= eps);
  bool bz = (std::fabs(rd.z) >= EPS_INV) && (std::fabs(slab.tmaxv.z - te) <= eps);
  uint32_t c = (bx ? 1u : 0u) + (by ? 1u : 0u) + (bz ? 1u : 0u);
  if (c >= 2u) return ExitFaceRes{0u, 1u};
  uint32_t axis = 0;
  if (c == 1u) {
    if (bx) axis = 0;
    if (by) axis = 1;
    if (bz) axis = 2;
  } else {
    float bestv = slab.tmaxv.x;
    axis = 0;
    if (slab.tmaxv.y < bestv) { bestv = slab.tmaxv.y; axis = 1; }
    if (slab.tmaxv.z < bestv) { bestv = slab.tmaxv.z; axis = 2; }
  }
  if (axis == 0) return ExitFaceRes{rd.x > 0.0f ? 0u : 1u, 0u};
  if (axis == 1) return ExitFaceRes{rd.y > 0.0f ? 2u : 3u, 0u};
  return ExitFaceRes{rd.z > 0.0f ? 4u : 5u, 0u};
}
// descend_leaf_sparse, chunk_trace.wgsl:296-358
static inline LeafState descend_leaf_sparse(const Ctx& C, v3 p_in, uint32_t node_base, uint32_t start_idx, v3 start_bmin,
                                            float start_size) {
  uint32_t idx = start_idx;
  v3 bmin = start_bmin;
  float size = start_size;
  uint32_t max_d = chunk_max_depth(C);
  for (uint32_t d = 0; d < 32u; ++d) {
    vx_node n = ld_node(C, (uint64_t)node_base + idx);
    if (n.child_base == LEAF_U32) return LeafState{true, idx, bmin, size, n.material, false, 0u, V3s(0.0f), 0.0f};
    if (d >= max_d) return LeafState{true, idx, bmin, size, MAT_AIR, false, 0u, V3s(0.0f), 0.0f};
    float half = size * 0.5f;
    v3 mid = bmin + V3s(half);
    float e = 1e-6f * size;
    uint32_t hx = p_in.x > mid.x + e ? 1u : 0u, hy = p_in.y > mid.y + e ? 1u : 0u, hz = p_in.z > mid.z + e ? 1u : 0u;
    uint32_t ci = hx | (hy << 1u) | (hz << 2u);
    v3 child_bmin = bmin + V3(hx ? half : 0.0f, hy ? half : 0.0f, hz ? half : 0.0f);
    uint32_t bit = 1u << ci;
    if ((n.child_mask & bit) == 0u)
      return LeafState{false, INVALID_U32, child_bmin, half, MAT_AIR, true, idx, bmin, size};
    idx = n.child_base + child_rank(n.child_mask, ci);
    bmin = child_bmin;
    size = half;
  }
  return LeafState{false, INVALID_U32, start_bmin, start_size, MAT_AIR, false, INVALID_U32, V3s(0.0f), 0.0f};
}

// Macro DDA, chunk_trace.wgsl:368-587
struct MacroDDAState {
  bool enabled;
  int32_t mx, my, mz, step_x, step_y, step_z;
  float tMaxX, tMaxY, tMaxZ, tDeltaX, tDeltaY, tDeltaZ, cell;
};
// macro_dda_init, chunk_trace.wgsl:397-475
static inline MacroDDAState macro_dda_init(v3 ro, v3 rd, v3 inv, float tcur, v3 root_bmin, float root_size, uint32_t macro_base) {
  MacroDDAState s;
  s.enabled = false;
  s.mx = s.my = s.mz = 0;
  s.step_x = s.step_y = s.step_z = 0;
  s.tMaxX = s.tMaxY = s.tMaxZ = BIG_F32;
  s.tDeltaX = s.tDeltaY = s.tDeltaZ = BIG_F32;
  s.cell = 0.0f;
  if (macro_base == INVALID_U32) return s;
  float cell = root_size / (float)MACRO_DIM;
  v3 p = ro + rd * tcur;
  v3 lp = p - root_bmin;
  if (lp.x < 0.0f || lp.y < 0.0f || lp.z < 0.0f || lp.x >= root_size || lp.y >= root_size || lp.z >= root_size) return s;
  int32_t mxi = f2i(std::floor(lp.x / cell)), myi = f2i(std::floor(lp.y / cell)), mzi = f2i(std::floor(lp.z / cell));
  if (mxi < 0 || myi < 0 || mzi < 0 || mxi >= (int32_t)MACRO_DIM || myi >= (int32_t)MACRO_DIM || mzi >= (int32_t)MACRO_DIM) return s;
  s.enabled = true;
  s.mx = mxi;
  s.my = myi;
  s.mz = mzi;
  s.step_x = rd.x > 0.0f ? 1 : -1;
  s.step_y = rd.y > 0.0f ? 1 : -1;
  s.step_z = rd.z > 0.0f ? 1 : -1;
  s.tDeltaX = std::fabs(cell * inv.x);
  s.tDeltaY = std::fabs(cell * inv.y);
  s.tDeltaZ = std::fabs(cell * inv.z);
  s.cell = cell;
  if (std::fabs(rd.x) >= EPS_INV) {
    int32_t nix = mxi + (rd.x > 0.0f ? 1 : 0);
    float bx = root_bmin.x + (float)nix * cell;
    s.tMaxX = tcur + (bx - p.x) * inv.x;
  }
  if (std::fabs(rd.y) >= EPS_INV) {
    int32_t niy = myi + (rd.y > 0.0f ? 1 : 0);
    float by = root_bmin.y + (float)niy * cell;
    s.tMaxY = tcur + (by - p.y) * inv.y;
  }
  if (std::fabs(rd.z) >= EPS_INV) {
    int32_t niz = mzi + (rd.z > 0.0f ? 1 : 0);
    float bz = root_bmin.z + (float)niz * cell;
    s.tMaxZ = tcur + (bz - p.z) * inv.z;
  }
  return s;
}
// macro_dda_t_exit, chunk_trace.wgsl:532-534
static inline float macro_dda_t_exit(const MacroDDAState& s) { return mn(s.tMaxX, mn(s.tMaxY, s.tMaxZ)); }
// macro_dda_step, chunk_trace.wgsl:536-559
static inline void macro_dda_step(MacroDDAState& s) {
  float tNext = macro_dda_t_exit(s);
  float epsTie = 1e-6f * mx(1.0f, std::fabs(tNext));
  if (std::fabs(s.tMaxX - tNext) <= epsTie) { s.mx += s.step_x; s.tMaxX += s.tDeltaX; }
  if (std::fabs(s.tMaxY - tNext) <= epsTie) { s.my += s.step_y; s.tMaxY += s.tDeltaY; }
  if (std::fabs(s.tMaxZ - tNext) <= epsTie) { s.mz += s.step_z; s.tMaxZ += s.tDeltaZ; }
  if (s.mx < 0 || s.my < 0 || s.mz < 0 || s.mx >= (int32_t)MACRO_DIM || s.my >= (int32_t)MACRO_DIM || s.mz >= (int32_t)MACRO_DIM)
    s.enabled = false;
}
// macro_dda_sync, chunk_trace.wgsl:561-568
static inline void macro_dda_sync(MacroDDAState& s, float tcur) {
  for (uint32_t i = 0; i < 32u; ++i) {
    if (!s.enabled) break;
    if (tcur < macro_dda_t_exit(s)) break;
    macro_dda_step(s);
  }
}
struct MacroStep { bool valid, empty; float t_exit; };
// macro_dda_current, chunk_trace.wgsl:570-587
static inline MacroStep macro_dda_current(const Ctx& C, const MacroDDAState& s, uint32_t macro_base) {
  MacroStep o;
  o.valid = s.enabled;
  o.empty = false;
  o.t_exit = macro_dda_t_exit(s);
  if (!s.enabled || macro_base == INVALID_U32) { o.valid = false; return o; }
  uint32_t bit = macro_bit_index((uint32_t)s.mx, (uint32_t)s.my, (uint32_t)s.mz);
  o.empty = !macro_test(C, macro_base, bit);
  return o;
}

// colinfo_entry_u16 / colinfo_decode, chunk_trace.wgsl:1284-1308
static inline uint32_t colinfo_entry_u16(const Ctx& C, uint32_t colinfo_base, uint32_t lx, uint32_t lz) {
  uint32_t idx = lz * 64u + lx;
  uint32_t word = ld_colinfo(C, (uint64_t)colinfo_base + (idx >> 1u));
  if ((idx & 1u) == 0u) return word & 0xFFFFu;
  return (word >> 16u) & 0xFFFFu;
}
struct ColInfo { bool valid; uint32_t y_vox, mat; };
static inline ColInfo colinfo_decode(uint32_t e16) {
  uint32_t y8 = e16 & 0xFFu, mat8 = (e16 >> 8u) & 0xFFu;
  if (y8 == 255u) return ColInfo{false, 0u, 0u};
  return ColInfo{true, y8, mat8};
}
// chunk_heightfield_early_skip, chunk_trace.wgsl:1310-1352
static inline bool chunk_heightfield_early_skip(const Ctx& C, v3 ro, v3 rd, float t_enter, float t_exit,
                                                const vx_chunk_meta& ch, float vs) {
  if (ch.colinfo_base == INVALID_U32) return false;
  if (rd.y >= 0.0f) return false;
  float y0 = ro.y + rd.y * t_enter, y1 = ro.y + rd.y * t_exit;
  float seg_y_min = mn(y0, y1);
  v3 root_bmin = V3((float)ch.origin[0], (float)ch.origin[1], (float)ch.origin[2]) * vs;
  float seg_len = mx(t_exit - t_enter, 0.0f);
  float dt = seg_len * (1.0f / 3.0f);
  bool found = false;
  for (uint32_t i = 0; i < 4u; ++i) {
    float t = t_enter + (float)i * dt;
    v3 p = ro + rd * t;
    float lx_f = (p.x - root_bmin.x) / vs, lz_f = (p.z - root_bmin.z) / vs;
    int32_t lx = clampi(f2i(std::floor(lx_f)), 0, 63), lz = clampi(f2i(std::floor(lz_f)), 0, 63);
    ColInfo ci = colinfo_decode(colinfo_entry_u16(C, ch.colinfo_base, (uint32_t)lx, (uint32_t)lz));
    if (!ci.valid) continue;
    float top_y = root_bmin.y + ((float)ci.y_vox + 1.0f) * vs;
    if (seg_y_min <= top_y + 1e-3f * vs) { found = true; break; }
  }
  return !found;
}
// probe_grass_columns_xz_dda, chunk_trace.wgsl:1358-1504
static inline GrassHit probe_grass_columns_xz_dda(const Ctx& C, v3 ro, v3 rd, v3 inv, float t0_in, float t1_in, v3 root_bmin,
                                                  i3 org, float vs, uint32_t colinfo_base, float time_s, float strength,
                                                  int32_t* hit_cell) {
  float t0 = t0_in, t1 = t1_in;
  if (t1 <= t0) return GrassHit{false, BIG_F32, V3s(0.0f)};
  float eps = 1e-4f * vs;
  float t = t0 + eps;
  if (t > t1) t = t0;
  v3 p = ro + rd * t;
  float lx_f = (p.x - root_bmin.x) / vs, lz_f = (p.z - root_bmin.z) / vs;
  int32_t lx = clampi(f2i(std::floor(lx_f)), 0, 63), lz = clampi(f2i(std::floor(lz_f)), 0, 63);
  int32_t stepX = rd.x > 0.0f ? 1 : -1, stepZ = rd.z > 0.0f ? 1 : -1;
  float tMaxX = BIG_F32, tDeltaX = BIG_F32;
  if (std::fabs(rd.x) >= EPS_INV) {
    float nextX = root_bmin.x + (float)(stepX > 0 ? lx + 1 : lx) * vs;
    tMaxX = t + (nextX - p.x) * inv.x;
    tDeltaX = vs * std::fabs(inv.x);
  }
  float tMaxZ = BIG_F32, tDeltaZ = BIG_F32;
  if (std::fabs(rd.z) >= EPS_INV) {
    float nextZ = root_bmin.z + (float)(stepZ > 0 ? lz + 1 : lz) * vs;
    tMaxZ = t + (nextZ - p.z) * inv.z;
    tDeltaZ = vs * std::fabs(inv.z);
  }
  for (uint32_t iter = 0; iter < 64u; ++iter) {
    if (t > t1) break;
    float tNext = mn(mn(tMaxX, tMaxZ), t1);
    ColInfo ci = colinfo_decode(colinfo_entry_u16(C, colinfo_base, (uint32_t)lx, (uint32_t)lz));
    if (ci.valid && ci.mat == MAT_GRASS) {
      int32_t wx = org.x + lx, wy = org.y + (int32_t)ci.y_vox, wz = org.z + lz;
      v3 cell_bmin = V3((float)wx, (float)wy, (float)wz) * vs;
      v3 id = V3((float)wx, (float)wy, (float)wz);
      float slab_y0 = cell_bmin.y + vs;
      float slab_y1 = slab_y0 + (GRASS_LAYER_HEIGHT_VOX * vs);
      float yA = ro.y + rd.y * t, yB = ro.y + rd.y * tNext;
      float sy0 = mn(yA, yB), sy1 = mx(yA, yB);
      if (!(sy1 < slab_y0 || sy0 > slab_y1)) {
        GrassHit gh = try_grass_slab_hit(C, ro, rd, t, tNext, cell_bmin, id, vs, time_s, strength);
        if (gh.hit) {
          hit_cell[0] = wx;
          hit_cell[1] = wy;
          hit_cell[2] = wz;
          return gh;
        }
      }
    }
    if (tNext >= t1) break;
    float tStep = mn(tMaxX, tMaxZ);
    float epsTie = 1e-6f * mx(1.0f, std::fabs(tStep));
    if (std::fabs(tMaxX - tStep) <= epsTie) { lx += stepX; tMaxX += tDeltaX; }
    if (std::fabs(tMaxZ - tStep) <= epsTie) { lz += stepZ; tMaxZ += tDeltaZ; }
    t = tStep + eps;
    if (lx < 0 || lx > 63 || lz < 0 || lz > 63) break;
    p = ro + rd * t;
  }
  return GrassHit{false, BIG_F32, V3s(0.0f)};
}

// Hit record helpers (identity is not part of the reference's HitGeom).
static inline HitGeom make_hit(uint32_t mat, float t, v3 n, v3 root_bmin, float root_size, const vx_chunk_meta& ch) {
  HitGeom h = miss_hitgeom();
  h.hit = 1u;
  h.t = t;
  h.mat = mat;
  h.n = n;
  h.root_bmin = root_bmin;
  h.root_size = root_size;
  h.node_base = ch.node_base;
  h.macro_base = ch.macro_base;
  return h;
}
static inline void set_leaf_identity(const Ctx& C, HitGeom& h, const vx_chunk_meta& ch, uint32_t idx_local) {
  uint32_t key = C.nodes[(uint64_t)ch.node_base + idx_local].key;  // not an algorithmic load
  uint32_t lvl = node_level(key);
  int32_t span = (int32_t)(C.cam->chunk_size >> lvl);
  h.node = idx_local;
  h.vox[0] = ch.origin[0] + (int32_t)node_x(key) * span;
  h.vox[1] = ch.origin[1] + (int32_t)node_y(key) * span;
  h.vox[2] = ch.origin[2] + (int32_t)node_z(key) * span;
}

// trace_chunk_rope_interval_nomacro, chunk_trace.wgsl:669-961
static inline ChunkTraceResult trace_chunk_rope_interval_nomacro(const Ctx& C, v3 ro, v3 rd, v3 inv, const vx_chunk_meta& ch,
                                                                 float t_enter, float t_exit, bool use_anchor, uint32_t anchor_key) {
  float vs = vp_x(C);
  v3 root_bmin = V3((float)ch.origin[0], (float)ch.origin[1], (float)ch.origin[2]) * vs;
  float root_size = (float)C.cam->chunk_size * vs;
  float eps_step = 1e-4f * vs;
  float tcur = mx(t_enter, 0.0f) + eps_step;
  AnchorNode anchor_node = anchor_none();
  if (use_anchor) anchor_node = anchor_from_key(C, ch.node_base, root_bmin, root_size, anchor_key);
  float grass_probe_max_leaf = 8.0f * vs;
  i3 org{ch.origin[0], ch.origin[1], ch.origin[2]};
  float time_s = C.cam->voxel_params[1], strength = C.cam->voxel_params[2];
  bool have_leaf = false;
  LeafState leaf;
  std::memset(&leaf, 0, sizeof leaf);
  uint32_t MAX_ITERS = 128u + C.cam->chunk_size * 2u;
  auto adopt_anchor = [&](const LeafState& lf) {
    AnchorNode na = anchor_from_leaf(C, lf, ch.node_base);
    if (na.valid && (!anchor_node.valid || na.key != anchor_node.key)) anchor_node = na;
  };
  for (uint32_t it = 0; it < MAX_ITERS; ++it) {
    if (tcur > t_exit) break;
    v3 p = ro + tcur * rd;
    v3 pq = p + ray_eps_vec(rd, 1e-4f * vs);
    if (!have_leaf || !point_in_cube(pq, leaf.bmin, leaf.size)) {
      if (anchor_node.valid && point_in_cube(pq, anchor_node.bmin, anchor_node.size))
        leaf = descend_leaf_sparse(C, pq, ch.node_base, anchor_node.idx, anchor_node.bmin, anchor_node.size);
      else
        leaf = descend_leaf_sparse(C, pq, ch.node_base, 0u, root_bmin, root_size);
      adopt_anchor(leaf);
      have_leaf = true;
    }
    CubeSlab slab = cube_slab_inv(ro, inv, leaf.bmin, leaf.size);
    float t_leave = slab.t_exit;
    if (leaf.mat == MAT_AIR) {
      uint32_t lod_probe = grass_lod_from_t(tcur);
      float limit = grass_probe_max_leaf;
      if (lod_probe == 1u) limit = 4.0f * vp_x(C);
      else if (lod_probe == 2u) limit = 2.0f * vp_x(C);
      if (!C.geom_only && leaf.size <= limit) {
        float t0p = mx(t_enter, tcur - eps_step);
        float t1p = mn(t_leave, t_exit);
        if (t1p >= t0p) {
          int32_t cell[3];
          GrassHit gh = probe_grass_columns_xz_dda(C, ro, rd, inv, t0p, t1p, root_bmin, org, vs, ch.colinfo_base, time_s,
                                                   strength, cell);
          if (gh.hit) {
            HitGeom o = make_hit(MAT_GRASS, gh.t, gh.n, root_bmin, root_size, ch);
            o.vox[0] = cell[0];
            o.vox[1] = cell[1];
            o.vox[2] = cell[2];
            return ChunkTraceResult{o, anchor_node.valid, anchor_node.key};
          }
        }
      }
      ExitFaceRes ex = exit_face_from_slab_safe(rd, slab);
      tcur = mx(t_leave, tcur) + eps_step;
      if (tcur > t_exit) break;
      if (ex.amb != 0u) { have_leaf = false; continue; }
      uint32_t face = ex.face;
      v3 p_next = ro + tcur * rd;
      v3 pq_next = p_next + ray_eps_vec(rd, 1e-4f * vs);
      if (leaf.has_node) {
        uint32_t nidx = rope_next_local(C, ch.node_base, leaf.idx_local, face);
        if (nidx == INVALID_U32) { have_leaf = false; continue; }
        uint32_t nk = ld_node(C, (uint64_t)ch.node_base + nidx).key;
        v4 c = node_cube_from_key(root_bmin, root_size, nk);
        leaf = descend_leaf_sparse(C, pq_next, ch.node_base, nidx, V3(c.x, c.y, c.z), c.w);
        adopt_anchor(leaf);
        have_leaf = true;
        continue;
      }
      if (leaf.has_anchor) {
        if (point_in_cube(pq_next, leaf.anchor_bmin, leaf.anchor_size)) {
          leaf = descend_leaf_sparse(C, pq_next, ch.node_base, leaf.anchor_idx, leaf.anchor_bmin, leaf.anchor_size);
          have_leaf = true;
          continue;
        }
        uint32_t nidx2 = rope_next_local(C, ch.node_base, leaf.anchor_idx, face);
        if (nidx2 == INVALID_U32) { have_leaf = false; continue; }
        uint32_t nk2 = ld_node(C, (uint64_t)ch.node_base + nidx2).key;
        v4 c2 = node_cube_from_key(root_bmin, root_size, nk2);
        leaf = descend_leaf_sparse(C, pq_next, ch.node_base, nidx2, V3(c2.x, c2.y, c2.z), c2.w);
        adopt_anchor(leaf);
        have_leaf = true;
        continue;
      }
      have_leaf = false;
      continue;
    }
    if (leaf.mat == MAT_LEAF && !C.geom_only) {
      float tmin_iter = mx(t_enter, tcur - eps_step);
      LeafCubeHit h2 = leaf_displaced_cube_hit(C, ro, rd, leaf.bmin, leaf.size, time_s, strength, tmin_iter, t_exit);
      if (h2.hit) {
        HitGeom o = make_hit(MAT_LEAF, h2.t, h2.n, root_bmin, root_size, ch);
        set_leaf_identity(C, o, ch, leaf.idx_local);
        return ChunkTraceResult{o, anchor_node.valid, anchor_node.key};
      }
      tcur = mx(t_leave, tcur) + eps_step;
      have_leaf = false;
      continue;
    }
    float tmin_iter = mx(t_enter, tcur - eps_step);
    BoxHit bh = cube_hit_normal_from_slab(rd, slab, tmin_iter, t_exit);
    if (bh.hit) {
      if (leaf.mat == MAT_GRASS && !C.geom_only) {
        v3 hp = ro + bh.t * rd;
        GrassCell cell = pick_grass_cell_in_chunk_with_bias(hp, bh.n, root_bmin, org, vs, (int32_t)C.cam->chunk_size);
        int32_t cid[3] = {f2i(cell.id_vox.x), f2i(cell.id_vox.y), f2i(cell.id_vox.z)};
        if (ch.colinfo_base != INVALID_U32) {
          int32_t lx = clampi(f2i(std::floor((hp.x - root_bmin.x) / vs)), 0, 63);
          int32_t lz = clampi(f2i(std::floor((hp.z - root_bmin.z) / vs)), 0, 63);
          ColInfo ci = colinfo_decode(colinfo_entry_u16(C, ch.colinfo_base, (uint32_t)lx, (uint32_t)lz));
          if (ci.valid && ci.mat == MAT_GRASS) {
            int32_t wx = org.x + lx, wy = org.y + (int32_t)ci.y_vox, wz = org.z + lz;
            cell.bmin_m = V3((float)wx, (float)wy, (float)wz) * vs;
            cell.id_vox = V3((float)wx, (float)wy, (float)wz);
            cid[0] = wx;
            cid[1] = wy;
            cid[2] = wz;
          }
        }
        float probe_pad = (GRASS_LAYER_HEIGHT_VOX + 1.0f) * vs;
        float tmin_probe = mx(t_enter, bh.t - 0.01f * vs);
        float tmax_probe = mn(bh.t + probe_pad, t_exit + probe_pad);
        GrassHit gh = try_grass_slab_hit(C, ro, rd, tmin_probe, tmax_probe, cell.bmin_m, cell.id_vox, vs, time_s, strength);
        if (gh.hit) {
          HitGeom o = make_hit(MAT_GRASS, gh.t, gh.n, root_bmin, root_size, ch);
          o.vox[0] = cid[0];
          o.vox[1] = cid[1];
          o.vox[2] = cid[2];
          return ChunkTraceResult{o, anchor_node.valid, anchor_node.key};
        }
      }
      HitGeom o = make_hit(leaf.mat, bh.t, bh.n, root_bmin, root_size, ch);
      set_leaf_identity(C, o, ch, leaf.idx_local);
      return ChunkTraceResult{o, anchor_node.valid, anchor_node.key};
    }
    tcur = mx(t_leave, tcur) + eps_step;
    have_leaf = false;
  }
  return ChunkTraceResult{miss_hitgeom(), anchor_node.valid, anchor_node.key};
}

// trace_chunk_interval_stream_macro, chunk_trace.wgsl:592-667
static inline ChunkTraceResult trace_chunk_interval_stream_macro(const Ctx& C, v3 ro, v3 rd, v3 inv, const vx_chunk_meta& ch,
                                                                 float t0_in, float t1_in, bool use_anchor, uint32_t anchor_key_in) {
  float vs = vp_x(C);
  v3 root_bmin = V3((float)ch.origin[0], (float)ch.origin[1], (float)ch.origin[2]) * vs;
  float root_size = (float)C.cam->chunk_size * vs;
  float eps_step = 1e-4f * vs;
  float tcur = mx(t0_in, 0.0f) + eps_step;
  float t_end = t1_in;
  uint32_t anchor_key = anchor_key_in;
  bool anchor_valid = use_anchor;
  if (ch.macro_base == INVALID_U32)
    return trace_chunk_rope_interval_nomacro(C, ro, rd, inv, ch, tcur, t_end, anchor_valid, anchor_key);
  MacroDDAState ms = macro_dda_init(ro, rd, inv, tcur, root_bmin, root_size, ch.macro_base);
  if (!ms.enabled) return trace_chunk_rope_interval_nomacro(C, ro, rd, inv, ch, tcur, t_end, anchor_valid, anchor_key);
  for (uint32_t i = 0; i < 128u; ++i) {
    if (tcur >= t_end) break;
    macro_dda_sync(ms, tcur);
    MacroStep st = macro_dda_current(C, ms, ch.macro_base);
    if (!st.valid) break;
    float t_cell_exit = mn(st.t_exit, t_end);
    if (st.empty) {
      tcur = t_cell_exit + eps_step;
      if (ms.enabled && tcur >= st.t_exit) macro_dda_step(ms);
      continue;
    }
    ChunkTraceResult h = trace_chunk_rope_interval_nomacro(C, ro, rd, inv, ch, tcur, t_cell_exit, anchor_valid, anchor_key);
    if (h.hit.hit != 0u) return h;
    anchor_valid = h.anchor_valid;
    anchor_key = h.anchor_key;
    tcur = t_cell_exit + eps_step;
    if (ms.enabled && tcur >= st.t_exit) macro_dda_step(ms);
  }
  return ChunkTraceResult{miss_hitgeom(), anchor_valid, anchor_key};
}

// Workgroup candidate list, common.wgsl:422-424.
struct TileList {
  uint32_t count;  // atomic counter value (may exceed MAX_TILE_CHUNKS)
  uint32_t slots[MAX_TILE_CHUNKS];
  float enter[MAX_TILE_CHUNKS];
  uint32_t cached;  // WG_TILE_COUNT_CACHED
};
// tile_add_candidate, chunk_trace.wgsl:1007-1026
static inline void tile_add_candidate(const Ctx& C, TileList& L, uint32_t slot, float cell_enter) {
  if (slot < C.cam->chunk_count) {
    vx_chunk_meta ch = ld_chunk(C, slot);
    if (ch.macro_empty != 0u) return;
  }
  uint32_t count = L.count;
  for (uint32_t i = 0; i < count && i < MAX_TILE_CHUNKS; ++i) {
    if (L.slots[i] == slot) {
      L.enter[i] = mn(L.enter[i], cell_enter);
      return;
    }
  }
  uint32_t idx = L.count++;
  if (idx < MAX_TILE_CHUNKS) {
    L.slots[idx] = slot;
    L.enter[idx] = cell_enter;
  }
}
// tile_append_candidates_for_ray, chunk_trace.wgsl:1028-1151
static inline void tile_append_candidates_for_ray(const Ctx& C, TileList& L, v3 ro, v3 rd, float t_min, float t_max) {
  if (C.cam->chunk_count == 0u) return;
  float vs = vp_x(C);
  float csm = (float)C.cam->chunk_size * vs;
  const int32_t* go = C.cam->grid_origin_chunk;
  const uint32_t* gd = C.cam->grid_dims;
  v3 gbmin = V3((float)go[0], (float)go[1], (float)go[2]) * csm;
  v3 gbmax = gbmin + V3((float)gd[0], (float)gd[1], (float)gd[2]) * csm;
  v2 rtg = intersect_aabb(ro, rd, gbmin, gbmax);
  float t_enter = mx(mx(rtg.x, 0.0f), t_min);
  float t_exit = mn(mn(rtg.y, FOG_MAX_DIST), t_max);
  if (t_exit < t_enter) return;
  float nudge_p = PRIMARY_NUDGE_VOXEL_FRAC * vs;
  float start_t = t_enter + nudge_p;
  v3 p0 = ro + start_t * rd;
  i3 c = chunk_coord_from_pos(p0, csm);
  int32_t nx = (int32_t)gd[0], ny = (int32_t)gd[1], nz = (int32_t)gd[2];
  int32_t lcx = c.x - go[0], lcy = c.y - go[1], lcz = c.z - go[2];
  if (lcx < 0 || lcy < 0 || lcz < 0 || lcx >= nx || lcy >= ny || lcz >= nz) return;
  int32_t stride_y = nx, stride_z = nx * ny;
  int32_t idx_i = (lcz * ny + lcy) * nx + lcx;
  float t_local = 0.0f;
  float t_exit_local = mx(t_exit - start_t, 0.0f);
  v3 inv = safe_inv3(rd);
  int32_t step_x = rd.x > 0.0f ? 1 : -1, step_y = rd.y > 0.0f ? 1 : -1, step_z = rd.z > 0.0f ? 1 : -1;
  float bx = rd.x > 0.0f ? (float)(c.x + 1) * csm : (float)c.x * csm;
  float by = rd.y > 0.0f ? (float)(c.y + 1) * csm : (float)c.y * csm;
  float bz = rd.z > 0.0f ? (float)(c.z + 1) * csm : (float)c.z * csm;
  float tMaxX = (bx - p0.x) * inv.x, tMaxY = (by - p0.y) * inv.y, tMaxZ = (bz - p0.z) * inv.z;
  float tDeltaX = std::fabs(csm * inv.x), tDeltaY = std::fabs(csm * inv.y), tDeltaZ = std::fabs(csm * inv.z);
  if (std::fabs(rd.x) < EPS_INV) tMaxX = BIG_F32;
  if (std::fabs(rd.y) < EPS_INV) tMaxY = BIG_F32;
  if (std::fabs(rd.z) < EPS_INV) tMaxZ = BIG_F32;
  int32_t didx_x = rd.x > 0.0f ? 1 : -1, didx_y = rd.y > 0.0f ? stride_y : -stride_y, didx_z = rd.z > 0.0f ? stride_z : -stride_z;
  int32_t rem_x = 0, rem_y = 0, rem_z = 0;
  if (std::fabs(rd.x) >= EPS_INV) rem_x = step_x > 0 ? (nx - 1 - lcx) : lcx;
  if (std::fabs(rd.y) >= EPS_INV) rem_y = step_y > 0 ? (ny - 1 - lcy) : lcy;
  if (std::fabs(rd.z) >= EPS_INV) rem_z = step_z > 0 ? (nz - 1 - lcz) : lcz;
  uint32_t msteps = (uint32_t)(rem_x + rem_y + rem_z) + 1u;
  if (msteps > 512u) msteps = 512u;
  for (uint32_t s = 0; s < msteps; ++s) {
    if (t_local > t_exit_local) break;
    uint32_t slot = ld_grid(C, (uint32_t)idx_i);
    if (slot != INVALID_U32 && slot < C.cam->chunk_count) tile_add_candidate(C, L, slot, start_t + t_local);
    float tNext = mn(tMaxX, mn(tMaxY, tMaxZ));
    float epsTie = 1e-6f * mx(1.0f, std::fabs(tNext));
    if (std::fabs(tMaxX - tNext) <= epsTie) { lcx += step_x; idx_i += didx_x; tMaxX += tDeltaX; }
    if (std::fabs(tMaxY - tNext) <= epsTie) { lcy += step_y; idx_i += didx_y; tMaxY += tDeltaY; }
    if (std::fabs(tMaxZ - tNext) <= epsTie) { lcz += step_z; idx_i += didx_z; tMaxZ += tDeltaZ; }
    t_local = tNext;
    if (lcx < 0 || lcy < 0 || lcz < 0 || lcx >= nx || lcy >= ny || lcz >= nz) break;
  }
}
// tile_sort_candidates_by_enter, chunk_trace.wgsl:1153-1175
static inline void tile_sort_candidates_by_enter(TileList& L, uint32_t count) {
  if (count <= 1u) return;
  uint32_t limit = count < PRIMARY_MAX_TILE_CHUNKS ? count : PRIMARY_MAX_TILE_CHUNKS;
  if (limit <= 1u) return;
  for (uint32_t i = 0; i + 1u < limit; ++i) {
    uint32_t mi = i;
    float mt = L.enter[i];
    for (uint32_t j = i + 1u; j < limit; ++j)
      if (L.enter[j] < mt) { mt = L.enter[j]; mi = j; }
    if (mi != i) {
      uint32_t ts = L.slots[i];
      float tt = L.enter[i];
      L.slots[i] = L.slots[mi];
      L.enter[i] = L.enter[mi];
      L.slots[mi] = ts;
      L.enter[mi] = tt;
    }
  }
}

struct VoxTraceResult { bool in_grid; HitGeom best; float t_exit; bool anchor_valid; uint32_t anchor_key; i3 anchor_chunk; };
// chunk_coords_neighbor, chunk_trace.wgsl:1000-1005
static inline bool chunk_coords_neighbor(i3 a, i3 b) {
  return std::abs(a.x - b.x) <= 1 && std::abs(a.y - b.y) <= 1 && std::abs(a.z - b.z) <= 1;
}
// trace_scene_voxels_candidates, chunk_trace.wgsl:1177-1265
static inline VoxTraceResult trace_scene_voxels_candidates(const Ctx& C, const TileList& L, v3 ro, v3 rd, float t_min, float t_max,
                                                           bool anchor_valid_in, i3 anchor_chunk_in, uint32_t anchor_key_in,
                                                           uint32_t candidate_count) {
  VoxTraceResult out{false, miss_hitgeom(), 0.0f, false, INVALID_U32, i3{0, 0, 0}};
  if (C.cam->chunk_count == 0u) return out;
  float vs = vp_x(C);
  float csm = (float)C.cam->chunk_size * vs;
  const int32_t* go = C.cam->grid_origin_chunk;
  const uint32_t* gd = C.cam->grid_dims;
  v3 gbmin = V3((float)go[0], (float)go[1], (float)go[2]) * csm;
  v3 gbmax = gbmin + V3((float)gd[0], (float)gd[1], (float)gd[2]) * csm;
  v2 rtg = intersect_aabb(ro, rd, gbmin, gbmax);
  float t_enter = mx(mx(rtg.x, 0.0f), t_min);
  float t_exit = mn(mn(rtg.y, FOG_MAX_DIST), t_max);
  if (t_exit < t_enter) return out;
  HitGeom best = miss_hitgeom();
  bool bav = false;
  uint32_t bak = INVALID_U32;
  i3 bac{0, 0, 0};
  v3 inv = safe_inv3(rd);
  float root_size = (float)C.cam->chunk_size * vs;
  uint32_t maxc = candidate_count < MAX_TILE_CHUNKS ? candidate_count : MAX_TILE_CHUNKS;
  for (uint32_t i = 0; i < maxc; ++i) {
    uint32_t slot = L.slots[i];
    if (slot == INVALID_U32 || slot >= C.cam->chunk_count) continue;
    vx_chunk_meta ch = ld_chunk(C, slot);
    if (ch.macro_empty != 0u) continue;
    v3 root_bmin = V3((float)ch.origin[0], (float)ch.origin[1], (float)ch.origin[2]) * vs;
    v3 root_bmax = root_bmin + V3s(root_size);
    v2 rtc = intersect_aabb(ro, rd, root_bmin, root_bmax);
    float cell_enter = mx(rtc.x, t_enter);
    float cell_exit = mn(rtc.y, t_exit);
    if (cell_exit < cell_enter) continue;
    if (best.hit != 0u && cell_enter >= best.t) continue;
    if (chunk_heightfield_early_skip(C, ro, rd, cell_enter, cell_exit, ch, vs)) continue;
    int32_t cs = (int32_t)C.cam->chunk_size;
    i3 cc{ch.origin[0] / cs, ch.origin[1] / cs, ch.origin[2] / cs};
    bool use_anchor = anchor_valid_in && chunk_coords_neighbor(cc, anchor_chunk_in);
    uint32_t aku = use_anchor ? anchor_key_in : INVALID_U32;
    ChunkTraceResult h = trace_chunk_interval_stream_macro(C, ro, rd, inv, ch, cell_enter, cell_exit, use_anchor, aku);
    if (h.hit.hit != 0u && h.hit.t < best.t) {
      best = h.hit;
      best.slot = slot;
      bav = h.anchor_valid;
      bak = h.anchor_key;
      bac = cc;
    }
  }
  return VoxTraceResult{true, best, t_exit, bav, bak, bac};
}

// ===========================================================================
// ray/shadows.wgsl
// ===========================================================================
// trace_chunk_shadow_trans_interval, shadows.wgsl:4-82
static inline float trace_chunk_shadow_trans_interval(const Ctx& C, v3 ro, v3 rd, const vx_chunk_meta& ch, float t_enter, float t_exit) {
  float vs = vp_x(C);
  float nudge_s = 0.20f * vs;
  v3 root_bmin = V3((float)ch.origin[0], (float)ch.origin[1], (float)ch.origin[2]) * vs;
  float root_size = (float)C.cam->chunk_size * vs;
  float tcur = mx(t_enter, 0.0f) + nudge_s;
  v3 inv = safe_inv3(rd);
  float trans = 1.0f;
  for (uint32_t si = 0; si < VSM_STEPS; ++si) {
    if (tcur > t_exit) break;
    if (trans < MIN_TRANS) break;
    v3 p = ro + tcur * rd;
    v3 pq = p + ray_eps_vec(rd, 1e-4f * vp_x(C));
    LeafQuery q = query_leaf_at(C, pq, root_bmin, root_size, ch.node_base, ch.macro_base);
    CubeSlab slab = cube_slab_inv(ro, inv, q.bmin, q.size);
    float t_leave = slab.t_exit;
    if (q.mat != MAT_AIR) {
      if (q.mat == MAT_LEAF) {
        v3 center = q.bmin + V3s(0.5f * q.size);
        float d = length3(center - cam_pos(C));
        if (d < LEAF_LOD_DISP_END) {
          LeafCubeHit h2 = leaf_displaced_cube_hit(C, ro, rd, q.bmin, q.size, C.cam->voxel_params[1], C.cam->voxel_params[2],
                                                   tcur - nudge_s, t_exit);
          if (h2.hit) trans *= LEAF_LIGHT_TRANSMIT;
        } else {
          trans *= LEAF_LIGHT_TRANSMIT;
        }
        tcur = mx(t_leave, tcur) + nudge_s;
        continue;
      }
      if (q.mat == MAT_GRASS) {
        trans *= GRASS_LIGHT_TRANSMIT;
        tcur = mx(t_leave, tcur) + nudge_s;
        continue;
      }
      return 0.0f;
    }
    tcur = mx(t_leave, tcur) + nudge_s;
  }
  return trans;
}
// sun_transmittance_geom_only, shadows.wgsl:85-186
static inline float sun_transmittance_geom_only(const Ctx& C, v3 p, v3 sun_dir) {
  float vs = vp_x(C);
  float nudge_s = 0.18f * vs;
  float csm = (float)C.cam->chunk_size * vs;
  const int32_t* go = C.cam->grid_origin_chunk;
  const uint32_t* gd = C.cam->grid_dims;
  v3 gbmin = V3((float)go[0], (float)go[1], (float)go[2]) * csm;
  v3 gbmax = gbmin + V3((float)gd[0], (float)gd[1], (float)gd[2]) * csm;
  float bias = mx(SHADOW_BIAS, 0.50f * vs);
  v3 ro = p + sun_dir * bias;
  v3 rd = sun_dir;
  v2 rtg = intersect_aabb(ro, rd, gbmin, gbmax);
  float t_enter = mx(rtg.x, 0.0f), t_exit = rtg.y;
  if (t_exit < t_enter) return 1.0f;
  float start_t = t_enter + nudge_s;
  v3 p0 = ro + start_t * rd;
  float t_local = 0.0f;
  float t_exit_local = mx(t_exit - start_t, 0.0f);
  i3 c = chunk_coord_from_pos(p0, csm);
  int32_t cx = c.x, cy = c.y, cz = c.z;
  v3 inv = safe_inv3(rd);
  int32_t step_x = rd.x > 0.0f ? 1 : -1, step_y = rd.y > 0.0f ? 1 : -1, step_z = rd.z > 0.0f ? 1 : -1;
  float bx = rd.x > 0.0f ? (float)(cx + 1) * csm : (float)cx * csm;
  float by = rd.y > 0.0f ? (float)(cy + 1) * csm : (float)cy * csm;
  float bz = rd.z > 0.0f ? (float)(cz + 1) * csm : (float)cz * csm;
  float tMaxX = (bx - p0.x) * inv.x, tMaxY = (by - p0.y) * inv.y, tMaxZ = (bz - p0.z) * inv.z;
  float tDeltaX = std::fabs(csm * inv.x), tDeltaY = std::fabs(csm * inv.y), tDeltaZ = std::fabs(csm * inv.z);
  if (std::fabs(rd.x) < EPS_INV) tMaxX = BIG_F32;
  if (std::fabs(rd.y) < EPS_INV) tMaxY = BIG_F32;
  if (std::fabs(rd.z) < EPS_INV) tMaxZ = BIG_F32;
  float trans = 1.0f;
  uint32_t msteps = (gd[0] + gd[1] + gd[2]) * 6u + 8u;
  if (msteps > 512u) msteps = 512u;
  int32_t gx0 = go[0], gy0 = go[1], gz0 = go[2];
  int32_t gx1 = go[0] + (int32_t)gd[0], gy1 = go[1] + (int32_t)gd[1], gz1 = go[2] + (int32_t)gd[2];
  for (uint32_t s = 0; s < msteps; ++s) {
    if (t_local > t_exit_local) break;
    if (trans < MIN_TRANS) break;
    float tNextLocal = mn(tMaxX, mn(tMaxY, tMaxZ));
    uint32_t slot = grid_lookup_slot(C, cx, cy, cz);
    if (slot != INVALID_U32 && slot < C.cam->chunk_count) {
      vx_chunk_meta ch2 = ld_chunk(C, slot);
      float ce = start_t + t_local;
      float cx2 = start_t + mn(tNextLocal, t_exit_local);
      trans *= trace_chunk_shadow_trans_interval(C, ro, rd, ch2, ce, cx2);
      if (trans < MIN_TRANS) break;
    }
    if (tMaxX < tMaxY) {
      if (tMaxX < tMaxZ) { cx += step_x; t_local = tMaxX; tMaxX += tDeltaX; }
      else { cz += step_z; t_local = tMaxZ; tMaxZ += tDeltaZ; }
    } else {
      if (tMaxY < tMaxZ) { cy += step_y; t_local = tMaxY; tMaxY += tDeltaY; }
      else { cz += step_z; t_local = tMaxZ; tMaxZ += tDeltaZ; }
    }
    if (cx < gx0 || cy < gy0 || cz < gz0 || cx >= gx1 || cy >= gy1 || cz >= gz1) break;
  }
  return trans;
}

// ===========================================================================
// Ray queries (no reference kernel; the reference picks voxels on the CPU,
// app/app.rs:847-961).  First voxel-geometry hit of a ray within [0, tmax]:
// a chunk-grid DDA in ray order; in each resident, non-empty chunk the
// chunk's own AABB interval is traced with the primary pass's macro DDA and
// rope traversal (trace_chunk_interval_stream_macro) in geometry-only mode.
// ===========================================================================
static inline HitGeom query_ray(const Ctx& C, v3 ro, v3 rd, float tmax) {
  const float vs = vp_x(C);
  const float csm = (float)C.cam->chunk_size * vs;
  const int32_t* go = C.cam->grid_origin_chunk;
  const uint32_t* gd = C.cam->grid_dims;
  const v3 gbmin = V3((float)go[0], (float)go[1], (float)go[2]) * csm;
  const v3 gbmax = gbmin + V3((float)gd[0], (float)gd[1], (float)gd[2]) * csm;
  const v2 rtg = intersect_aabb(ro, rd, gbmin, gbmax);
  const float t_enter = mx(rtg.x, 0.0f), t_exit = mn(rtg.y, tmax);
  if (t_exit < t_enter) return miss_hitgeom();
  const v3 inv = safe_inv3(rd);
  const v3 p0 = ro + rd * (t_enter + 1e-3f * vs);
  i3 c = chunk_coord_from_pos(p0, csm);
  const int32_t sx = rd.x > 0.0f ? 1 : -1, sy = rd.y > 0.0f ? 1 : -1, sz = rd.z > 0.0f ? 1 : -1;
  float tMaxX = ((rd.x > 0.0f ? (float)(c.x + 1) : (float)c.x) * csm - ro.x) * inv.x;
  float tMaxY = ((rd.y > 0.0f ? (float)(c.y + 1) : (float)c.y) * csm - ro.y) * inv.y;
  float tMaxZ = ((rd.z > 0.0f ? (float)(c.z + 1) : (float)c.z) * csm - ro.z) * inv.z;
  const float tDX = std::fabs(csm * inv.x), tDY = std::fabs(csm * inv.y), tDZ = std::fabs(csm * inv.z);
  if (std::fabs(rd.x) < EPS_INV) tMaxX = BIG_F32;
  if (std::fabs(rd.y) < EPS_INV) tMaxY = BIG_F32;
  if (std::fabs(rd.z) < EPS_INV) tMaxZ = BIG_F32;
  const int32_t gx1 = go[0] + (int32_t)gd[0], gy1 = go[1] + (int32_t)gd[1], gz1 = go[2] + (int32_t)gd[2];
  const uint32_t msteps = gd[0] + gd[1] + gd[2] + 2u;
  float t_cell = t_enter;
  for (uint32_t s = 0; s < msteps; ++s) {
    if (t_cell > t_exit) break;
    if (c.x < go[0] || c.y < go[1] || c.z < go[2] || c.x >= gx1 || c.y >= gy1 || c.z >= gz1) break;
    const uint32_t slot = grid_lookup_slot(C, c.x, c.y, c.z);
    if (slot != INVALID_U32 && slot < C.cam->chunk_count) {
      const vx_chunk_meta ch = ld_chunk(C, slot);
      if (ch.macro_empty == 0u) {
        const v3 rb = V3((float)ch.origin[0], (float)ch.origin[1], (float)ch.origin[2]) * vs;
        const v2 rc = intersect_aabb(ro, rd, rb, rb + V3s(csm));
        const float a = mx(rc.x, t_enter), b = mn(rc.y, t_exit);
        if (b >= a) {
          ChunkTraceResult r = trace_chunk_interval_stream_macro(C, ro, rd, inv, ch, a, b, false, INVALID_U32);
          if (r.hit.hit != 0u) {
            r.hit.slot = slot;
            return r.hit;
          }
        }
      }
    }
    if (tMaxX < tMaxY) {
      if (tMaxX < tMaxZ) { c.x += sx; t_cell = tMaxX; tMaxX += tDX; }
      else { c.z += sz; t_cell = tMaxZ; tMaxZ += tDZ; }
    } else {
      if (tMaxY < tMaxZ) { c.y += sy; t_cell = tMaxY; tMaxY += tDY; }
      else { c.z += sz; t_cell = tMaxZ; tMaxZ += tDZ; }
    }
  }
  return miss_hitgeom();
}

// ===========================================================================
// ray/light.wgsl
// ===========================================================================
// hash3_i32, light.wgsl:18-23
static inline uint32_t hash3_i32(i3 p) {
  uint32_t ux = (uint32_t)p.x * 73856093u, uy = (uint32_t)p.y * 19349663u, uz = (uint32_t)p.z * 83492791u;
  return hash_u32(ux ^ uy ^ uz);
}
// rot_about_axis, light.wgsl:32-38
static inline v3 rot_about_axis(v3 v, v3 axis, float ang) {
  v3 a = normalize3(axis);
  float s = vx_sin(ang), c = vx_cos(ang);
  return v * c + cross3(a, v) * s + a * dot3(a, v) * (1.0f - c);
}
// sphere_dir_local, light.wgsl:48-56
static inline v3 sphere_dir_local(uint32_t i, uint32_t count) {
  float fi = (float)i;
  float count_f = mx(1.0f, (float)count);
  const float golden = 2.399963229728653f;
  float y = 1.0f - 2.0f * ((fi + 0.5f) / count_f);
  float r = std::sqrt(mx(0.0f, 1.0f - y * y));
  float phi = golden * fi;
  return V3(vx_cos(phi) * r, vx_sin(phi) * r, y);
}
// dda_hit_light, light.wgsl:65-121
static inline v4 dda_hit_light(const Ctx& C, v3 p0, v3 dir, uint32_t max_steps, float vs) {
  v3 gpos = p0 / mx(vs, 1e-6f);
  v3 fg = floor3(gpos);
  i3 cell{f2i(fg.x), f2i(fg.y), f2i(fg.z)};
  v3 d = dir;
  i3 step{d.x >= 0.0f ? 1 : -1, d.y >= 0.0f ? 1 : -1, d.z >= 0.0f ? 1 : -1};
  v3 nb = V3((float)(cell.x + (d.x >= 0.0f ? 1 : 0)), (float)(cell.y + (d.y >= 0.0f ? 1 : 0)), (float)(cell.z + (d.z >= 0.0f ? 1 : 0)));
  v3 inv = V3(std::fabs(d.x) > 1e-6f ? 1.0f / std::fabs(d.x) : 1e30f, std::fabs(d.y) > 1e-6f ? 1.0f / std::fabs(d.y) : 1e30f,
              std::fabs(d.z) > 1e-6f ? 1.0f / std::fabs(d.z) : 1e30f);
  v3 tMax = V3((nb.x - gpos.x) * inv.x, (nb.y - gpos.y) * inv.y, (nb.z - gpos.z) * inv.z);
  v3 tDelta = inv;
  for (uint32_t s = 0; s < max_steps; ++s) {
    v3 center = V3((float)cell.x + 0.5f, (float)cell.y + 0.5f, (float)cell.z + 0.5f) * vs;
    LeafQuery leaf = query_leaf_world(C, center);
    if (leaf.mat == MAT_LIGHT) return v4{center.x, center.y, center.z, 1.0f};
    if (leaf.mat != MAT_AIR) return v4{0.0f, 0.0f, 0.0f, 0.0f};
    if (tMax.x < tMax.y) {
      if (tMax.x < tMax.z) { cell.x += step.x; tMax.x += tDelta.x; }
      else { cell.z += step.z; tMax.z += tDelta.z; }
    } else {
      if (tMax.y < tMax.z) { cell.y += step.y; tMax.y += tDelta.y; }
      else { cell.z += step.z; tMax.z += tDelta.z; }
    }
  }
  return v4{0.0f, 0.0f, 0.0f, 0.0f};
}
// gather_voxel_lights, light.wgsl:127-211
static inline v3 gather_voxel_lights(const Ctx& C, v3 hp, v3 n, v3 root_bmin) {
  uint32_t hits = 0;
  float vs = vp_x(C);
  v3 p0 = hp + n * (0.75f * vs);
  // make_tbn, light.wgsl:25-30
  v3 up_ref = std::fabs(n.y) > 0.9f ? V3(1.0f, 0.0f, 0.0f) : V3(0.0f, 1.0f, 0.0f);
  v3 tt = normalize3(cross3(up_ref, n));
  v3 bb = normalize3(cross3(n, tt));
  v3 sv = floor3((hp - root_bmin) / mx(vs, 1e-6f));
  uint32_t h0 = hash3_i32(i3{f2i(sv.x), f2i(sv.y), f2i(sv.z)});
  uint32_t h1 = hash_u32(h0);
  uint32_t h2 = hash_u32(h1 ^ (C.cam->frame_index * 747796405u));
  float rot = 6.28318530718f * ((float)(h2 & 1023u) / 1024.0f);
  float soft_r = LIGHT_SOFT_RADIUS_VOX * vs, soft_r2 = soft_r * soft_r;
  float min_r = LIGHT_NEAR_CLAMP_VOX * vs, min_r2 = min_r * min_r;
  float range_m = LIGHT_RANGE_VOX * vs, range2 = range_m * range_m;
  v3 Le = 18.0f * V3(1.0f, 0.95f, 0.75f);
  v3 sum = V3s(0.0f);
  for (uint32_t i = 0; i < LIGHT_RAYS; ++i) {
    v3 sl = normalize3(sphere_dir_local(i, LIGHT_RAYS));
    v3 tv = tt * sl.x + bb * sl.y + n * sl.z;  // mat3(t, b, n) * v
    v3 ldir = normalize3(tv);
    ldir = normalize3(rot_about_axis(ldir, n, rot));
    v4 hit = dda_hit_light(C, p0, ldir, LIGHT_MAX_DIST_VOX, vs);
    if (hit.w > 0.5f) {
      hits += 1u;
      v3 pL = V3(hit.x, hit.y, hit.z);
      v3 L = pL - hp;
      float r2 = mx(dot3(L, L), min_r2);
      float r = std::sqrt(r2);
      v3 ldws = L / r;
      float ndl_raw = dot3(n, ldws);
      float ndl = clampf((ndl_raw + LIGHT_WRAP) / (1.0f + LIGHT_WRAP), 0.0f, 1.0f);
      float att = clampf(1.0f - (r2 / mx(range2, 1e-6f)), 0.0f, 1.0f);
      att = att * att;
      float falloff = 1.0f / (r2 + soft_r2);
      sum = sum + Le * (LIGHT_DIRECT_GAIN * ndl * falloff * att);
      if (r2 < 0.35f * range2) {
        float bounce = LIGHT_INDIRECT_GAIN * falloff * att;
        sum = sum + Le * (0.25f * bounce);
      }
      if (hits >= LIGHT_EARLY_HITS) break;
    }
  }
  float denom = mx(1.0f, (float)LIGHT_RAYS);
  return sum * (1.0f / denom);
}

}  // namespace vxo
