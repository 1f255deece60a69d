// vx_launch.h — host-side interface between the runtime (vx_runtime.hip) and
// the kernel launchers (vx_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "vxd_scene.h"

// The primary pass's load-counting form, one kernel per 8x8 tile: the
// reference algorithm's loads, or (executed) the loads the product kernel
// makes with its exact work skips.  Adds the array counts and the part byte
// totals into counters and writes each ray's counted bytes to ray_bytes when
// it is not null.
hipError_t vx_launch_primary_count(const vxd::Scene& S, const vx_camera* cam_dev, const vxd::FrameBuffers& fb,
                                   bool executed, unsigned long long* counters, float* ray_bytes, hipStream_t st);
hipError_t vx_launch_local_taa(const vxd::FrameBuffers& fb, hipStream_t st);
hipError_t vx_launch_godray(const vxd::Scene& S, const vx_camera* cam_dev, const vxd::FrameBuffers& fb, bool count,
                            unsigned long long* counters, hipStream_t st);
hipError_t vx_launch_composite(const vxd::Scene& S, const vx_camera* cam_dev, const vxd::FrameBuffers& fb, hipStream_t st);
hipError_t vx_launch_composite_taa(const vxd::Scene& S, const vx_camera* cam_dev, const vxd::FrameBuffers& fb, hipStream_t st);
// A batch of views: fused work-stealing primary and godray kernels, flattened
// streaming passes, and the end-of-frame ping flip, in reference pass order.
// profile: launch the primary kernel that raises the profile counters of the
// views whose camera has profile_flags & 1 (the other form has no such code).
hipError_t vx_launch_batch(const vxd::Scene& S, const vxd::Scene& S_godray, const vxd::Batch& B, uint32_t mask,
                           uint32_t persistent, bool profile, hipStream_t st);
// One view through the same work-stealing primary kernel: reset the tile
// cursors, then a persistent launch of `persistent` workgroups.
hipError_t vx_launch_ws_begin(const vxd::Batch& B, hipStream_t st);
hipError_t vx_launch_primary_ws(const vxd::Scene& S, const vxd::Batch& B, uint32_t persistent, bool profile,
                                hipStream_t st);
// Ray queries over n rays (device arrays) as one persistent launch with lane refill.
hipError_t vx_launch_query(const vxd::Scene& S, const vx_camera* cam_dev, const vx_ray* rays, vx_hit* hits, uint32_t n,
                           uint32_t* cursor, uint32_t persistent, uint32_t refill_below, hipStream_t st);
// Ray queries in a given order: order[k] is the k-th ray to trace (nullptr:
// input order).  Hits go to their rays' own positions.
hipError_t vx_launch_query_ordered(const vxd::Scene& S, const vx_camera* cam_dev, const vx_ray* rays, vx_hit* hits,
                                   uint32_t n, const uint32_t* order, uint32_t* cursor, uint32_t persistent,
                                   uint32_t refill_below, hipStream_t st);
// Sort keys of n rays (keys[i], vals[i] = i): direction bins (3 bits per axis)
// above the origin's Morton code over the chunk grid's box (5 bits per axis),
// kQuerySortBits bits in all.
constexpr uint32_t kQuerySortBits = 24;
hipError_t vx_launch_query_keys(const vx_camera* cam_dev, const vx_ray* rays, uint32_t n, uint32_t* keys, uint32_t* vals,
                                hipStream_t st);
// rocPRIM's radix sort of (key, value) pairs on bits [0, end_bit); with temp
// null it only sets *temp_bytes.
hipError_t vx_sort_pairs_u32(void* temp, size_t* temp_bytes, const uint32_t* keys_in, uint32_t* keys_out,
                             const uint32_t* vals_in, uint32_t* vals_out, uint32_t n, uint32_t end_bit, hipStream_t st);
// The ray queries' load-counting form: one ray per lane, the reference
// algorithm's loads or (executed) those k_query makes, added into counters as
// by vx_launch_primary_count; the hits are those of vx_launch_query.
hipError_t vx_launch_query_count(const vxd::Scene& S, const vx_camera* cam_dev, const vx_ray* rays, vx_hit* hits, uint32_t n,
                                 bool executed, unsigned long long* counters, hipStream_t st);
// Per chunk slot: 1 when its node range holds a LIGHT leaf, else 0 (n_slots entries).
hipError_t vx_launch_chunk_lights(const vxd::Scene& S, uint32_t* out, uint32_t n_slots, hipStream_t st);
// Clipmap texels of n rectangles (device array, at most max_texels texels each)
// from the world generator, stored at their storage positions in clip_height.
hipError_t vx_launch_clip_fill(const vx_clip_rect* rects, uint32_t n, uint32_t max_texels, const vx_world_params& world,
                               float* clip_height, hipStream_t st);
