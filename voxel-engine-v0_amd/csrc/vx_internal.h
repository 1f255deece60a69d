// vx_internal.h — host-side helpers shared by the library's translation units.
#pragma once

#include <cstdarg>
#include <cstddef>
#include <cstdint>

#include "voxhip.h"

static_assert(sizeof(vx_node) == 16, "NodeGpu is 16 bytes (gpu_types.rs:15-20)");
static_assert(sizeof(vx_node_ropes) == 32, "NodeRopesGpu is 32 bytes (gpu_types.rs:24-33)");
static_assert(sizeof(vx_chunk_meta) == 48, "ChunkMetaGpu is 48 bytes (gpu_types.rs:38-48)");
static_assert(sizeof(vx_camera) == 464, "CameraGpu is 464 bytes (gpu_types.rs:52-86)");
static_assert(sizeof(vx_clipmap) == 528, "ClipmapGpu is 528 bytes (gpu_types.rs:105-113)");
static_assert(offsetof(vx_camera, cam_pos) == 256, "CameraGpu.cam_pos @256");
static_assert(offsetof(vx_camera, chunk_size) == 320, "CameraGpu.chunk_size @320");
static_assert(offsetof(vx_camera, voxel_params) == 336, "CameraGpu.voxel_params @336");
static_assert(offsetof(vx_camera, grid_origin_chunk) == 352, "CameraGpu.grid_origin_chunk @352");
static_assert(offsetof(vx_camera, render_present_px) == 384, "CameraGpu.render_present_px @384");
static_assert(offsetof(vx_camera, profile_flags) == 400, "CameraGpu.profile_flags @400");
static_assert(offsetof(vx_clipmap, offset) == 272, "ClipmapGpu.offset @272");
static_assert(sizeof(vx_hit) == 32, "vx_hit is 32 bytes");
static_assert(sizeof(vx_ray) == 32, "vx_ray is 32 bytes");

// Records a message for vx_last_error() on this thread and returns `code`.
int vx_fail(int code, const char* fmt, ...);

// Context properties for the library's other translation units.
uint64_t vx_internal_node_capacity(const vx_ctx* ctx);
int vx_internal_device(const vx_ctx* ctx);

// The resident arenas of a context, for the library's device-side producers.
struct vx_internal_arenas {
  vx_node* nodes;
  vx_node_ropes* ropes;
  uint32_t* macro;
  uint32_t* colinfo;
  uint64_t node_cap;
  uint32_t chunk_cap;
  void* stream;  // the context's hipStream_t
  int device;
};
void vx_internal_get_arenas(const vx_ctx* ctx, vx_internal_arenas* out);
// Marks the node arena as changed (per-slot derived flags are refreshed before the next render).
void vx_internal_nodes_changed(vx_ctx* ctx);
// Raises (+1) or lowers (-1) the context's count of live child objects
// (views, batches, servers, chunk builders); vx_destroy refuses while it is
// above zero.
void vx_internal_child(vx_ctx* ctx, int delta);

// The device chunk builder over chunks at any positions of resident material
// and ground arrays: batch chunk i reads position index[i] (host array; null:
// position i).  solid (host, may be null) receives whether each chunk holds a
// solid voxel.  The emit step (vx_chunk_builder_emit) reads the same positions.
struct vx_chunk_builder;
int vx_internal_builder_count(vx_chunk_builder* b, uint32_t n, const uint8_t* material, const int32_t* ground,
                              const uint32_t* index, const int32_t* chunk_oy, int32_t dirt_depth, uint32_t* node_counts,
                              uint8_t* solid);
// dst[off[i]] = val[i] for device arrays off / val (distinct offsets), on the context stream.
int vx_internal_scatter_u8(vx_ctx* ctx, uint8_t* dst, const uint64_t* d_off, const uint8_t* d_val, uint32_t n);
// Points an existing batch at another set of views of its context (the same
// checks as vx_batch_create), reusing its device arrays when they are large
// enough.  The previous render of the batch must have completed.
int vx_internal_batch_rebind(vx_batch* b, vx_view* const* views, uint32_t n);
