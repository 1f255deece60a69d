// vxd_scene.h — device-side view of the resident world and the bound camera,
// plus the per-load algorithmic byte counters used by the counting variant
// of the kernels.
#pragma once
#include "voxhip.h"
#include "vxd_math.h"

namespace vxd {

// Resident world, as uploaded through the C ABI (render/gpu_types.rs formats).
struct Scene {
  const vx_chunk_meta* __restrict__ chunks;
  const uint4* __restrict__ nodes;    // vx_node as one 16-byte vector
  const uint4* __restrict__ ropes;    // vx_node_ropes as two 16-byte vectors
  const uint32_t* __restrict__ grid;
  const uint32_t* __restrict__ macro;  // macro occupancy words (or rope words, godray compat)
  const uint32_t* __restrict__ colinfo;
  const vx_clipmap* __restrict__ clip;
  const float* __restrict__ clip_height;
  uint64_t n_chunks, n_nodes, n_grid, n_macro, n_colinfo, n_clip;
  // Per chunk slot: nonzero when the slot's nodes hold a LIGHT leaf (derived
  // from the node arena and the chunk table, never uploaded).  Null: unknown,
  // so every local-light gather runs in full.
  const uint32_t* __restrict__ chunk_light;
  // The primary kernel's deferred grass tests: one region of kDeferEntries x
  // 64 entries per workgroup (entry e of lane l at [e * 64 + l]), for at most
  // n_defer_regions workgroups.
  uint4* __restrict__ defer;
  uint32_t n_defer_regions;
};

// Grass layer tests the primary traversal records per lane and per attempt
// at a chunk before the wavefront runs them together.
constexpr uint32_t kDeferEntries = 32;
// Workgroups per CU the primary kernel's deferral regions are sized for: its
// register budget (128 VGPRs, 4 waves per SIMD) and LDS allow at most 16.
constexpr uint32_t kDeferRegionsPerCU = 16;

// Parts of the primary pass that loads are charged to (VX_COUNT_PARTS):
// the tile's candidate gather, the voxel traversal, grass (column probes and
// blade refinement), the sun shadow, voxel shading (ambient occlusion from the
// macro words), the local-light gather, the heightfield (clipmap), and texture
// traffic (histories read, outputs written).
enum CountPart : uint32_t {
  kPartGather = 0, kPartTrace, kPartGrass, kPartShadow, kPartShade, kPartLight, kPartHeightfield, kPartTextures,
  kNumParts
};
// Algorithmic loads by array (element counts; bytes = count * element size),
// and the same loads' bytes by part of the pass.
struct Counts {
  uint32_t grid, meta, macro, node, rope, colinfo, clip, tex_rd, tex_wr;
  uint32_t part_bytes[kNumParts];
  uint32_t part;  // the part loads are charged to now
};
constexpr int kNumCounts = 9;

}  // namespace vxd

namespace vxd {

// Per-view textures of one frame, as linear row-major arrays (render res for
// primary/local TAA/godray, present res for composite/composite TAA).
struct FrameBuffers {
  uint32_t rw, rh, pw, ph;
  const float4* __restrict__ primary_hist_in;
  float4* __restrict__ primary_hist_out;
  const float* __restrict__ shadow_in;
  float* __restrict__ shadow_out;
  float4* __restrict__ color;
  float* __restrict__ depth;
  float4* __restrict__ local;
  const float4* __restrict__ local_hist_in;
  float4* __restrict__ local_hist_out;  // written by local TAA, read by composite
  const float4* __restrict__ godray_hist;
  float4* __restrict__ godray_out;      // written by godray, read by composite
  float4* __restrict__ pre_taa;
  const float4* __restrict__ out_hist_in;
  float4* __restrict__ out_hist_out;
  float4* __restrict__ output;
  vx_hit* __restrict__ hits;
  // The view's five profile counters (ray_main.wgsl:32-47), raised by the
  // primary pass when the camera's profile_flags & 1.
  uint32_t* __restrict__ prof;
};

// A fused multi-view launch: one table entry per view (both history sets, so
// the table stays valid while the views' ping-pong flips frame to frame),
// prefix offsets of the views' work items, and the work-stealing cursors.
struct BatchView {
  FrameBuffers fb[2];                       // by the view's current ping
  const vx_camera* cam;
  uint32_t tiles_x, tiles, tile_base;       // 8x8 tiles at render resolution
  uint32_t rblk_x, rblks, rblk_base;        // 16x16 blocks at render resolution
  uint32_t pblk_x, pblks, pblk_base;        // 16x16 blocks at present resolution
  // Primary execution tiles of the other shapes (16x4, 32x2): per row and
  // the view's first index in the batch's execution tile space of that shape.
  uint32_t xt_x[2], xt_base[2];
  uint32_t pad;
};
constexpr uint32_t kBatchBands = 8;         // one tile band per XCD
// Ray queries a fused launch traces after the views' primary tiles: device
// memory owned by the batch, rewritten from a pinned copy before a launch
// when they change.  The cursor is cleared by the launch's first kernel.
struct BatchRays {
  vx_camera cam;                            // the scene camera of the queries
  const vx_ray* rays;
  vx_hit* hits;
  const uint32_t* order;                    // the rays in the order to trace them (null: input order)
  uint32_t n;
  uint32_t refill_below;                    // lanes refill when fewer than this are busy
  uint32_t cursor;                          // next ray
  uint32_t pad;
};
// Primary execution tile shapes: 8x8 (the candidate list's own tile), 16x4
// and 32x2.  Whatever the shape, each pixel traces with the candidate list of
// its 8x8 tile; shapes other than 8x8 take the lists from a pre-pass.
constexpr uint32_t kTileShapes = 3;
constexpr uint32_t kListWords = 32;         // per 8x8 tile: the count, then up to 24 chunk slots
struct Batch {
  const BatchView* __restrict__ views;
  uint32_t* __restrict__ ping;              // per view, flipped by the end-of-frame kernel
  uint32_t* __restrict__ cursor;            // [pass][band]: next tile of each band (primary, godray)
  uint32_t n, tiles, rblks, pblks;
  uint32_t grab;                            // tiles taken per atomic
  BatchRays* rays;                          // null: the launch traces no ray queries
  uint32_t shape;                           // primary execution tile shape (0: 8x8, 1: 16x4, 2: 32x2)
  uint32_t xtiles[2];                       // execution tiles of shapes 1 and 2 over all views
  uint32_t* __restrict__ lists;             // shapes 1, 2: [8x8 tile][kListWords] from the pre-pass
};

// Workgroup-shared primary candidate list (common.wgsl:422-424).
constexpr uint32_t kMaxTileChunks = 192;
// Blade cache words per workgroup: 3 blades (the most any grass LOD draws) x
// 11 floats x 64 lanes.  With the tile list this keeps a workgroup's LDS
// under 10 KiB, so 16 one-wavefront workgroups fit a CU's 160 KiB.
constexpr uint32_t kGrassBladeCache = 3u * 11u * 64u;
struct TileLDS {
  uint32_t count;
  uint32_t cached;
  uint32_t slots[kMaxTileChunks];
  float enter[kMaxTileChunks];
  float grass_taper[12];  // blade taper per (lod, segment), see grass_taper_entry
  float grass_blades[kGrassBladeCache];  // per-lane blade cache, see grass_cache_store
  uint32_t prof[5];  // the tile's profile counter increments, flushed to FrameBuffers::prof per tile
};

}  // namespace vxd
