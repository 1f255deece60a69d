/*
 * voxhip.h — C ABI of libvoxhip, the MI355X (gfx950) implementation of the
 * voxel-engine-v0 primary / local-TAA / godray / composite passes.
 *
 * Two groups of entry points:
 *   1. The renderer boundary.  It replaces the reference's `render::Renderer`
 *      (src/render/state/mod.rs:34-712) and the WGSL binding table
 *      (src/shaders/common.wgsl:403-409, src/shaders/ray_main.wgsl:20-59).
 *      Inputs are the reference's GPU byte formats, unchanged
 *      (src/render/gpu_types.rs:15-138).
 *   2. The CPU producers of those inputs.  They replace the chunk builder
 *      (src/svo/builder.rs:846-986), the chunk meta/grid packing
 *      (src/streaming/manager/slots.rs:231-251, grid.rs:15-72), the camera
 *      packing (src/app/app.rs:411-511) and the clipmap fill
 *      (src/clipmap.rs:96-122,182-577).
 *
 * Conventions: every function returns an int status (VX_OK = 0, negative on
 * error) and never unwinds across the ABI.  vx_last_error() returns a
 * thread-local message for the last failing call on that thread.  Host
 * pointers are read before the call returns (the caller keeps ownership).
 * A context is externally synchronised: calls on one context (and its views)
 * must not overlap.
 */
#ifndef VOXHIP_H
#define VOXHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Status codes                                                             */
/* ------------------------------------------------------------------------ */
#define VX_OK            0
#define VX_E_INVALID    -1  /* bad argument (null pointer, size, enum)       */
#define VX_E_CAPACITY   -2  /* write past a capacity fixed at vx_create      */
#define VX_E_DEVICE     -3  /* HIP runtime error                             */
#define VX_E_NOMEM      -4  /* device or host allocation failed              */
#define VX_E_NODEVICE   -5  /* no usable GPU                                 */
#define VX_E_BUSY       -6  /* rejected by admission; vx_destroy: live children */
#define VX_E_AGAIN      -7  /* serving: nothing finished yet                 */

#define VX_INVALID_U32  0xFFFFFFFFu

/* Materials (src/world/materials.rs:1-7, common.wgsl:37-43). */
#define VX_MAT_AIR   0u
#define VX_MAT_GRASS 1u
#define VX_MAT_DIRT  2u
#define VX_MAT_STONE 3u
#define VX_MAT_WOOD  4u
#define VX_MAT_LEAF  5u
#define VX_MAT_LIGHT 6u

#define VX_CHUNK_SIZE        64u     /* app/config.rs:5                      */
#define VX_MACRO_WORDS       16u     /* common.wgsl:417                      */
#define VX_COLINFO_WORDS     2048u   /* svo/builder.rs:747                   */
#define VX_CLIP_LEVELS_MAX   16u     /* common.wgsl:241, app/config.rs:40    */
#define VX_MAX_CHUNK_NODES   299593u /* 64^3+32^3+16^3+8^3+4^3+2^3+1         */

/* ------------------------------------------------------------------------ */
/* Byte formats (identical to src/render/gpu_types.rs)                      */
/* ------------------------------------------------------------------------ */

/* NodeGpu, gpu_types.rs:15-20.  child_base == 0xFFFFFFFF marks a leaf.
 * key = x | y<<6 | z<<12 | level<<18 (svo/builder_svo.rs:22-35). */
typedef struct vx_node {
    uint32_t child_base;
    uint32_t child_mask;
    uint32_t material;
    uint32_t key;
} vx_node;

/* NodeRopesGpu, gpu_types.rs:24-33.  Chunk-local neighbour indices. */
typedef struct vx_node_ropes {
    uint32_t px, nx, py, ny, pz, nz;
    uint32_t pad0, pad1;
} vx_node_ropes;

/* ChunkMetaGpu, gpu_types.rs:38-48. */
typedef struct vx_chunk_meta {
    int32_t  origin[4];      /* chunk origin in voxels */
    uint32_t node_base;      /* arena offset (elements) of this chunk's nodes/ropes */
    uint32_t node_count;
    uint32_t macro_base;     /* = slot * 16  */
    uint32_t colinfo_base;   /* = slot * 2048 */
    uint32_t macro_empty;    /* 1 iff all 16 macro words are zero */
    uint32_t pad0, pad1, pad2;
} vx_chunk_meta;

/* CameraGpu, gpu_types.rs:52-86 (464 bytes, column-major matrices). */
typedef struct vx_camera {
    float    view_inv[16];
    float    proj_inv[16];
    float    view_proj[16];
    float    prev_view_proj[16];
    float    cam_pos[4];
    float    ray00[4];
    float    ray_dx[4];
    float    ray_dy[4];
    uint32_t chunk_size;
    uint32_t chunk_count;
    uint32_t max_steps;
    uint32_t frame_index;
    float    voxel_params[4];      /* [voxel_size_m, time_s, wind_strength, fog_density] */
    int32_t  grid_origin_chunk[4];
    uint32_t grid_dims[4];
    uint32_t render_present_px[4]; /* render w, render h, present w, present h */
    uint32_t profile_flags;
    uint32_t pad_profile[3];
    uint32_t pad0[4];
    uint32_t pad1[4];
    uint32_t pad2[4];
} vx_camera;

/* ClipmapGpu, gpu_types.rs:105-113 (528 bytes). */
typedef struct vx_clipmap {
    uint32_t levels;
    uint32_t res;
    float    base_cell_m;
    float    pad0;
    float    level[16][4];   /* origin_x_m, origin_z_m, cell_m, 0 */
    uint32_t offset[16][4];  /* toroidal offsets off_x, off_z, 0, 0 */
} vx_clipmap;

/* Per-pixel hit identity (not present in the reference, whose HitGeom carries
 * no node index: ray/chunk_trace.wgsl:15-32).  Written when requested. */
typedef struct vx_hit {
    uint32_t slot;     /* chunk slot of the hit, VX_INVALID_U32 if none          */
    uint32_t node;     /* chunk-local leaf node index, VX_INVALID_U32 for grass  */
    int32_t  voxel[3]; /* world voxel of the hit leaf's min corner / grass cell  */
    uint32_t mat;      /* material, VX_MAT_AIR for miss                          */
    float    t;        /* hit distance in metres (BIG for miss)                  */
    uint32_t kind;     /* 0 sky, 1 voxel, 2 heightfield                           */
} vx_hit;

/* ------------------------------------------------------------------------ */
/* 1. Renderer boundary                                                     */
/* ------------------------------------------------------------------------ */

typedef struct vx_ctx  vx_ctx;
typedef struct vx_view vx_view;

/* Capacities (render/state/buffers.rs:79-127).  Zero means the reference's
 * default: 67,108,864 nodes, 12,996 chunks/grid cells, 16 clip levels x 256^2. */
typedef struct vx_caps {
    uint64_t node_capacity;     /* elements of the node and rope arenas        */
    uint32_t chunk_capacity;    /* chunk meta slots (macro = x16, colinfo = x2048) */
    uint32_t grid_capacity;     /* chunk grid cells                             */
    uint32_t clip_levels;       /* clip height array layers                     */
    uint32_t clip_res;          /* clip height array resolution                 */
    int32_t  device;            /* HIP device ordinal                           */
    uint32_t flags;             /* VX_CTX_* */
} vx_caps;

/* By default the godray pass reads macro occupancy through the rope buffer, as
 * the reference does: its scene bind group puts the ropes at the binding the
 * shader declares for macro_occ (render/state/bindgroups.rs:142-151 vs
 * common.wgsl:407-409).  VX_CTX_GODRAY_MACRO_FIX makes it read the macro words. */
#define VX_CTX_GODRAY_MACRO_FIX 1u

/* Renderer::new (render/state/mod.rs:147-268): allocates the resident world.
 * vx_destroy frees it (the Renderer's drop).  Teardown order: every object
 * created on the context (views, batches, servers, chunk builders) must be
 * destroyed first, because each of them launches on the context's stream and
 * reads its arenas (a server's device thread keeps doing so until
 * vx_server_destroy joins it).  While any of them is alive vx_destroy frees
 * nothing and returns VX_E_BUSY; the context stays valid. */
int vx_create(const vx_caps* caps, vx_ctx** out);
int vx_destroy(vx_ctx* ctx);

/* Arenas addressed by vx_upload_regions (render/state/mod.rs:514-582). */
#define VX_ARENA_NODES   0
#define VX_ARENA_ROPES   1
#define VX_ARENA_META    2
#define VX_ARENA_MACRO   3
#define VX_ARENA_COLINFO 4

typedef struct vx_region {
    uint64_t    offset;    /* element offset in the arena                      */
    uint64_t    count;     /* elements                                         */
    const void* data;      /* host pointer, count * element size bytes         */
} vx_region;

/* Renderer::apply_chunk_uploads (render/state/mod.rs:514-582).  Regions past
 * the capacity fail with VX_E_CAPACITY (the reference skips them silently). */
int vx_upload_regions(vx_ctx* ctx, int arena, const vx_region* regions, uint32_t n);

/* Renderer::write_chunk_grid (render/state/mod.rs:310-317).  The grid origin and
 * dims travel in vx_camera, as in the reference; here only the cell values. */
int vx_write_grid(vx_ctx* ctx, const uint32_t* grid, uint32_t cells);
/* Copies count elements at an element offset of an arena back to host memory
 * (no reference counterpart: for tests and tools). */
int vx_read_region(vx_ctx* ctx, int arena, uint64_t offset, uint64_t count, void* dst);

/* Renderer::write_clipmap_updates (render/state/mod.rs:586-634).  Patches are
 * rectangles of one level in storage (toroidal) texel coordinates. */
typedef struct vx_clip_patch {
    uint32_t level, x, y, w, h;
    const float* data;     /* w*h floats, row-major */
} vx_clip_patch;
int vx_write_clipmap(vx_ctx* ctx, const vx_clipmap* params, const vx_clip_patch* patches, uint32_t n);
/* Renderer::write_clipmap_updates with the heights generated on the device:
 * one kernel computes every texel of the rectangles (the same heights as
 * vx_clipmap_fill_rects, bit for bit) straight into the resident clip array,
 * then the uniform is written; stream-ordered before the next vx_render, like
 * vx_write_clipmap.  At most VX_CLIP_RECTS_MAX rectangles; `world` is the seeded
 * world generator's parameters (section 2). */
struct vx_world_params;
struct vx_clip_rect;
int vx_write_clipmap_rects(vx_ctx* ctx, const vx_clipmap* params, const struct vx_world_params* world,
                           const struct vx_clip_rect* rects, uint32_t n);
/* Copies the resident clip heights (clip_levels * clip_res^2 floats, storage
 * order) to host memory; bytes must cover them. */
int vx_read_clipmap(vx_ctx* ctx, float* heights, size_t bytes);

/* A view: one camera stream with its own history (the reference's texture set,
 * render/state/textures.rs:122-285, and ping-pong, bindgroups.rs:305-459).
 * render = internal resolution of primary/local_taa/godray; present = output. */
int  vx_view_create(vx_ctx* ctx, uint32_t present_w, uint32_t present_h,
                    uint32_t render_w, uint32_t render_h, vx_view** out);
void vx_view_destroy(vx_view* view);
/* Renderer::resize_output (render/state/mod.rs:297-308): reallocates and clears history. */
int  vx_view_resize(vx_view* view, uint32_t present_w, uint32_t present_h,
                    uint32_t render_w, uint32_t render_h);
/* Clears all history textures to zero (a fresh Renderer). */
int  vx_view_reset(vx_view* view);

/* Renderer::write_camera (render/state/mod.rs:319-322). */
int vx_write_camera(vx_view* view, const vx_camera* cam);

/* Renderer::encode_compute (render/state/mod.rs:329-446).  pass_mask selects
 * the passes run this call, in reference order, against the current history
 * slot; VX_PASS_END_FRAME then flips the ping-pong (render/state/mod.rs:445),
 * so a frame may be split over several calls.  stream: a hipStream_t or NULL
 * (the context's stream).  Launches are stream-ordered; the call does not block. */
#define VX_PASS_PRIMARY        0x01u
#define VX_PASS_LOCAL_TAA      0x02u
#define VX_PASS_GODRAY         0x04u
#define VX_PASS_COMPOSITE      0x08u
#define VX_PASS_COMPOSITE_TAA  0x10u
#define VX_PASS_END_FRAME      0x80u   /* flip the history ping-pong after the passes */
#define VX_PASS_FRAME          0x9Fu   /* all five passes + end of frame */
#define VX_PASS_WRITE_HITS     0x100u  /* also write the vx_hit buffer in primary */
#define VX_PASS_COUNT          0x200u  /* run the load-counting variant of primary/godray */
/* With VX_PASS_COUNT: the primary pass counts the loads the product kernel
 * executes (its exact work skips on: light gathers with no light in reach,
 * air regions of the light rays) instead of the reference algorithm's. */
#define VX_PASS_COUNT_EXECUTED 0x400u
int vx_render(vx_view* view, uint32_t pass_mask, void* stream);

/* Output selectors for vx_read_output / vx_output_device_ptr.  Linear row-major
 * arrays; RGBA32F = 16 B/px, R32F = 4 B/px.  History selectors name the copy the
 * NEXT frame will read (i.e. the one this frame wrote). */
#define VX_OUT_COLOR          0  /* render res RGBA32F (primary colour)        */
#define VX_OUT_DEPTH          1  /* render res R32F                            */
#define VX_OUT_LOCAL          2  /* render res RGBA32F                         */
#define VX_OUT_PRIMARY_HIST   3  /* render res RGBA32F                         */
#define VX_OUT_SHADOW_HIST    4  /* render res R32F (dense, no row padding)    */
#define VX_OUT_LOCAL_HIST     5  /* render res RGBA32F                         */
#define VX_OUT_GODRAY         6  /* render res RGBA32F                         */
#define VX_OUT_PRE_TAA        7  /* present res RGBA32F (composite output)     */
#define VX_OUT_OUTPUT         8  /* present res RGBA32F (composite TAA output) */
#define VX_OUT_HITS           9  /* render res vx_hit                          */
/* render res R32F: per ray, the bytes of the reference algorithm's loads and
 * stores counted by the last VX_PASS_COUNT primary pass (the tile's candidate
 * gather shared over its 64 rays); no reference counterpart. */
#define VX_OUT_RAY_BYTES     10
/* Synchronous device-to-host copy of one output (waits for the view's stream). */
int vx_read_output(vx_view* view, int which, void* dst, size_t bytes);
/* Device pointer of one output, for zero-copy consumers on the same device. */
int vx_output_device_ptr(vx_view* view, int which, void** dptr, size_t* bytes);
/* Host-to-device write of a history texture (tests seed history with it). */
int vx_write_history(vx_view* view, int which, const void* src, size_t bytes);

/* Algorithmic load counters accumulated by VX_PASS_COUNT runs since the last
 * reset: elements loaded from [grid, meta, macro, node, rope, colinfo, clip]
 * and texture bytes [read, written]. */
int vx_view_counters(vx_view* view, uint64_t out[9], int reset);
/* The same counted loads' bytes by part of the primary pass (no reference
 * counterpart): [tile candidate gather, voxel traversal, grass (column probes
 * and blade refinement), sun shadow, voxel shading, local-light gather,
 * heightfield, textures (history reads and output writes)].  Accumulated with
 * vx_view_counters' counts and cleared by its reset. */
#define VX_COUNT_PARTS 8
int vx_view_count_parts(vx_view* view, uint64_t out[VX_COUNT_PARTS]);

/* The primary pass's profile counters (ray_main.wgsl:32-74), raised while the
 * camera's profile_flags & 1 by every primary pass of this view (vx_render
 * and vx_batch_render): [PROFILE_VOXEL (candidate chunks per voxel trace),
 * PROFILE_GRASS (grass refinement hits), PROFILE_HDR (shading calls, 2 per
 * voxel hit and 1 per heightfield hit), PROFILE_FOG (fog weight
 * max(1, 256 * clamp(t / 120, 0, 1))), PROFILE_SHADOW (2 per traced sun
 * shadow, 1 per reused one)], u32 and wrapping like the reference's atomics.
 * Waits for the device, copies them out and, when reset != 0, zeroes them.
 * Replaces Renderer::read_primary_profile_counts_blocking and
 * reset_primary_profile_counts (render/state/mod.rs:448-491). */
int vx_view_profile(vx_view* view, uint32_t out[5], int reset);

/* Per-pass GPU milliseconds of the last vx_render that ran them
 * (render/state/mod.rs:649-709): [primary, local_taa, godray, composite, composite_taa].
 * Waits for the view's stream. */
int vx_view_timings(vx_view* view, float ms[5]);

/* Ray queries (no reference counterpart beyond app/app.rs:847-961's CPU pick):
 * the first voxel hit of each ray against the resident world within [0, tmax],
 * voxel geometry only (leaf cubes; no grass blades, no leaf displacement).
 * scene_cam (host) supplies the grid origin/dims, chunk count and voxel size.
 * Rays and hits are DEVICE pointers; the call is stream-ordered and returns
 * without waiting.  All n rays run as ONE persistent work-stealing launch in
 * which lanes whose ray finished take new rays (lane refill).  Hit: kind 1,
 * slot, leaf node (chunk-local), voxel (world voxel coords of the leaf's
 * minimum corner), mat, t.  Miss: kind 0, slot/node 0xFFFFFFFF.
 * A context has one query camera buffer and one ray cursor: queries issued on
 * one stream are ordered by it, but a query on another stream must not start
 * before the previous one has finished (they would race on both). */
typedef struct vx_ray {
    float ox, oy, oz, tmax;   /* origin (metres), max distance (metres) */
    float dx, dy, dz, pad;    /* direction (unit) */
} vx_ray;
int vx_query_rays(vx_ctx* ctx, const vx_camera* scene_cam, const vx_ray* rays_dev,
                  uint32_t n, vx_hit* hits_dev, void* stream);
/* GPU milliseconds of the last vx_query_rays (waits for it). */
int vx_query_timing(vx_ctx* ctx, float* ms);
/* Algorithmic load counts of a query over rays_dev[0, n) (no reference
 * counterpart: the measurement form of vx_query_rays, as vx_view_counters is
 * of the primary pass).  Each ray is traced alone on one lane with per-lane
 * counters: the reference algorithm's loads (executed = 0) or the loads
 * vx_query_rays makes (executed = 1).  out[9] as vx_view_counters: elements
 * loaded from [grid, meta, macro, node, rope, colinfo, clip], then bytes of
 * [ray records read, hit records written]; parts (may be NULL) as
 * vx_view_count_parts.  hits_dev receives the hits vx_query_rays writes.
 * Synchronous: waits for the context's stream before and after.  It uses the
 * context's query camera buffer, so it must not overlap a vx_query_rays
 * issued on another stream. */
int vx_query_count(vx_ctx* ctx, const vx_camera* scene_cam, const vx_ray* rays_dev, uint32_t n,
                   vx_hit* hits_dev, int executed, uint64_t out[9], uint64_t parts[VX_COUNT_PARTS]);
typedef struct vx_query_config {
    uint32_t waves_per_simd;  /* persistent workgroups = CUs * 4 SIMDs * this (default 8) */
    uint32_t refill_below;    /* refill idle lanes when fewer lanes than this are busy (default 48) */
    uint32_t sort_rays;       /* 1 (default): a call of 2^20 rays or more traces them sorted by direction,
                                 then origin (a device radix sort), so a wavefront's rays walk the same
                                 chunks; 0: input order.  The hits are the same either way. */
    uint32_t pad;
} vx_query_config;
int vx_query_set_config(vx_ctx* ctx, const vx_query_config* cfg);
int vx_query_get_config(const vx_ctx* ctx, vx_query_config* cfg);
/* Autotuner: runs the query for waves_per_simd in {4, 6, 8} x refill_below in
 * {16, 32, 48, 64}, keeps the fastest (returned in *chosen, may be NULL).
 * sort_rays keeps its current value. */
int vx_query_tune(vx_ctx* ctx, const vx_camera* scene_cam, const vx_ray* rays_dev, uint32_t n,
                  vx_hit* hits_dev, vx_query_config* chosen);

/* Batches: many views of one context rendered together (the serving and
 * batching path; no reference counterpart: the reference renders one view per
 * encode_compute).  Each pass runs as ONE launch over all views: primary and
 * godray as persistent work-stealing kernels over the union of the views' 8x8
 * tiles (largest views first, one tile band per XCD, stealing across bands),
 * the streaming passes as one flattened grid.  Results are identical to
 * rendering each view with vx_render.  Each view keeps its own camera
 * (vx_write_camera) and history; VX_PASS_END_FRAME flips every view's history.
 * A batch refers to its views' buffers: resizing a view invalidates the batch
 * (vx_batch_render then fails with VX_E_INVALID). */
typedef struct vx_batch vx_batch;
typedef struct vx_batch_config {
    uint32_t waves_per_simd;  /* persistent workgroups = CUs * 4 SIMDs * this (default 6) */
    uint32_t grab;            /* tiles taken per work-stealing atomic (default 1) */
    uint32_t use_graph;       /* replay the pass sequence from a captured hipGraph (default 1) */
    uint32_t tile_shape;      /* primary execution tile: 0 8x8 (default), 1 16x4, 2 32x2; each
                                 pixel still traces with its 8x8 tile's candidate list, which
                                 shapes 1 and 2 build in a pre-pass (bit-identical outputs) */
} vx_batch_config;
int  vx_batch_create(vx_ctx* ctx, vx_view* const* views, uint32_t n, vx_batch** out);
void vx_batch_destroy(vx_batch* batch);
/* pass_mask as for vx_render, without VX_PASS_WRITE_HITS / VX_PASS_COUNT. */
int  vx_batch_render(vx_batch* batch, uint32_t pass_mask, void* stream);
/* GPU milliseconds of the last vx_batch_render (waits for it). */
int  vx_batch_timing(vx_batch* batch, float* ms);
int  vx_batch_set_config(vx_batch* batch, const vx_batch_config* cfg);
int  vx_batch_get_config(const vx_batch* batch, vx_batch_config* cfg);
/* Autotuner: times the passes in pass_mask (history is not advanced) for
 * tile_shape in {8x8, 16x4, 32x2} x waves_per_simd in {4, 8} x grab in {1, 2},
 * `iters` runs each, keeps the fastest configuration and returns it in
 * *chosen (may be NULL). */
int  vx_batch_tune(vx_batch* batch, uint32_t pass_mask, uint32_t iters, vx_batch_config* chosen);
/* No counterpart in the reference: ray batches fused into a batch's launch.
 * Each later vx_batch_render whose pass_mask holds VX_PASS_PRIMARY also
 * traces these n rays inside its persistent primary launch: workgroups take
 * 64-ray blocks (with the lane refill of vx_query_rays and the context's
 * query configuration) once every view tile has been taken, so the rays fill
 * the time the last tiles take.  hits_dev receives exactly what vx_query_rays
 * would write for the same scene_cam and rays.  The arrays stay the caller's
 * and must stay valid until the render has completed on its stream; a render
 * without a new call traces the same rays again.  n = 0 detaches them.  The
 * call waits until the last render that uploaded the previous rays has
 * started reading them. */
int  vx_batch_set_rays(vx_batch* batch, const vx_camera* scene_cam, const vx_ray* rays_dev, uint32_t n,
                       vx_hit* hits_dev);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc).  Renders up to VX_NOALLOC_MAX_VIEWS views of ctx as one fused
 * launch sequence, exactly as a vx_batch over the same views would
 * (vx_batch_render, pass_mask likewise without hits or counting), but with no
 * vx_batch object and without allocating: the launch's view table, ping words
 * and tile cursors live in a static device array (about 24 KB) that each GPU
 * keeps for the life of the process, resolved once per device by vx_create.
 * More views than VX_NOALLOC_MAX_VIEWS fail with VX_E_CAPACITY.  Every call on
 * one GPU uses the same array (across contexts): calls that overlap on a GPU,
 * from several threads or on several streams, race and must be serialised by
 * the caller.  The call waits for its table upload on `stream` (a hipStream_t
 * or NULL for the context's stream) and queues the passes without waiting. */
#define VX_NOALLOC_MAX_VIEWS 64u
int vx_render_views_noalloc(vx_ctx* ctx, vx_view* const* views, uint32_t n, uint32_t pass_mask, void* stream);

/* Blocks until all work queued on the context's stream is done. */
int vx_sync(vx_ctx* ctx);

const char* vx_last_error(void);
/* Library build string: compiler, target, flags. */
const char* vx_version(void);
/* 1 if a usable GPU is present, 0 otherwise (no error). */
int vx_device_count(void);

/* ------------------------------------------------------------------------ */
/* 2. CPU producers of the inputs                                           */
/* ------------------------------------------------------------------------ */

/* Chunk builder, svo/builder.rs:846-986 stages col tops -> colinfo -> prefix ->
 * macro bits -> min/max mip -> bottom-up SVO -> ropes.  material is the y-slab
 * layout idx = ly*4096 + lz*64 + lx (svo/builder.rs:599-658); ground is the
 * per-column ground height in world voxels (idx = lz*64 + lx), used only by the
 * stone-collapse rule (svo/builder_svo.rs:144-181).  nodes/ropes must hold
 * node_cap elements (VX_MAX_CHUNK_NODES always suffices). */
int vx_build_chunk(const uint8_t* material, const int32_t* ground, int32_t chunk_oy,
                   int32_t dirt_depth, vx_node* nodes, vx_node_ropes* ropes,
                   uint32_t node_cap, uint32_t* node_count,
                   uint32_t macro_words[16], uint32_t colinfo_words[2048]);

/* The chunk builder on the device: for a batch of chunks, the same node,
 * rope, macro and colinfo bytes as vx_build_chunk, built on the GPU straight
 * into the context's resident arenas (no reference counterpart: the reference
 * builds on CPU worker threads, svo/builder.rs:846-986).  Two steps:
 *   vx_chunk_builder_count: material (n * 64^3 bytes, y-slab layout, as for
 *     vx_build_chunk) and ground (n * 64^2 int32) are DEVICE arrays, chunk_oy a
 *     host array; writes each chunk's node count to node_counts (host).  The
 *     material must stay valid until the emit step returns.
 *   vx_chunk_builder_emit: writes the counted chunks at the given element
 *     offsets of the node and rope arenas and of the macro and colinfo words
 *     (the ChunkMetaGpu node_base / macro_base / colinfo_base the host then
 *     uploads).  Ranges past the context's capacities fail with VX_E_CAPACITY.
 * Both complete before they return.  The workspace holds about 450 KB of HBM
 * per chunk of max_chunks (at most VX_BUILD_CHUNKS_MAX). */
typedef struct vx_chunk_builder vx_chunk_builder;
#define VX_BUILD_CHUNKS_MAX 1024u
int  vx_chunk_builder_create(vx_ctx* ctx, uint32_t max_chunks, vx_chunk_builder** out);
void vx_chunk_builder_destroy(vx_chunk_builder* b);
int  vx_chunk_builder_count(vx_chunk_builder* b, uint32_t n, const uint8_t* material, const int32_t* ground,
                            const int32_t* chunk_oy, int32_t dirt_depth, uint32_t* node_counts);
int  vx_chunk_builder_emit(vx_chunk_builder* b, const uint64_t* node_base, const uint32_t* macro_base,
                           const uint32_t* colinfo_base);

/* The seeded synthetic world: 2-octave value-noise heightfield, stone/dirt/grass
 * layers and trees with leaf blobs.  Its definition is this library's (the
 * reference world generator, world/generator.rs, is not used). */
typedef struct vx_world_params {
    uint32_t seed;          /* 42 */
    int32_t  base_y;        /* 96 voxels */
    int32_t  amplitude;     /* 48 voxels */
    int32_t  dirt_depth;    /* 4: stone below ground-4, then dirt, grass top */
    float    tree_density;  /* 0.01 of columns */
    int32_t  trunk_min;     /* 5 */
    int32_t  trunk_max;     /* 12 */
    int32_t  leaf_radius;   /* 2 => 5^3 leaf blob */
    int32_t  wavelength0;   /* 64 voxels (octave 0); octave 1 is half */
} vx_world_params;

void    vx_world_default(vx_world_params* p);
int32_t vx_world_ground_height(const vx_world_params* p, int32_t wx, int32_t wz);
/* Fills one chunk's material (y-slab layout) and ground (lz*64+lx) arrays. */
int vx_world_fill_chunk(const vx_world_params* p, int32_t cx, int32_t cy, int32_t cz,
                        uint8_t* material, int32_t* ground);

/* A whole resident world built on the host: every chunk of the grid built with
 * vx_build_chunk (threads: 0 = hardware concurrency), packed into arenas in
 * grid order (x fastest, then y, then z), one slot per non-empty chunk,
 * node_base assigned contiguously (try_make_uploading, slots.rs:163-275). */
typedef struct vx_host_world vx_host_world;
int  vx_host_world_build(const vx_world_params* p, const int32_t grid_origin[3],
                         const uint32_t grid_dims[3], uint32_t threads, vx_host_world** out);
void vx_host_world_free(vx_host_world* w);
typedef struct vx_host_world_view {
    const vx_node*       nodes;    uint64_t node_count;
    const vx_node_ropes* ropes;
    const vx_chunk_meta* meta;     uint32_t chunk_count;
    const uint32_t*      macro_words;   /* chunk_count * 16   */
    const uint32_t*      colinfo_words; /* chunk_count * 2048 */
    const uint32_t*      grid;     uint32_t grid_cells;
    int32_t  grid_origin[3];
    uint32_t grid_dims[3];
} vx_host_world_view;
int vx_host_world_get(const vx_host_world* w, vx_host_world_view* out);
/* Applies single-voxel edits (world voxel coords, material) and rebuilds the
 * touched chunks in place on the host; returns the rebuilt slots.  Used by the
 * serving edit stream. */
int vx_host_world_edit(vx_host_world* w, const int32_t* xyz, const uint8_t* mats,
                       uint32_t n, uint32_t* rebuilt_slots, uint32_t* n_rebuilt);

/* Camera pose -> CameraGpu, app/app.rs:411-511 and app/camera.rs:95-111
 * (glam look_at_rh / perspective_rh, Halton(2,3) jitter of frame_index+1). */
typedef struct vx_pose {
    float pos[3];
    float yaw, pitch;       /* radians, camera.rs:96-100 */
    float fovy;             /* radians */
    float z_near, z_far;
} vx_pose;
#define VX_CAM_JITTER 1u
int vx_camera_build(const vx_pose* pose, uint32_t render_w, uint32_t render_h,
                    uint32_t present_w, uint32_t present_h, uint32_t frame_index,
                    float time_s, float wind_strength, float fog_density,
                    const int32_t grid_origin[3], const uint32_t grid_dims[3],
                    uint32_t chunk_count, const float* prev_view_proj /* 16 or NULL */,
                    uint32_t flags, vx_camera* out);

/* Full clipmap for a camera position, clipmap.rs:62-122 (snapped origins,
 * texel-centre heights) with toroidal offsets = origin cell mod res.
 * heights: levels*res*res floats, layer-major, row-major (storage coordinates). */
int vx_clipmap_build(const vx_world_params* p, float cam_x, float cam_z, uint32_t levels,
                     uint32_t res, float base_cell_m, vx_clipmap* out, float* heights);

/* Incremental clipmap: the state of Clipmap (clipmap.rs:44-60) and its per-frame
 * Clipmap::update (clipmap.rs:182-577).  Per level the state keeps the committed
 * origin cell and toroidal offset (storage = (logical + offset) mod res).  An
 * update snaps every level's origin to the camera, then, if the frame updates
 * (a level is new, the camera is within VX_CLIP_EDGE_GUARD texels of a level's
 * committed edge, or a level moved and min_update_interval_s has passed since
 * the last update, clipmap.rs:276-284), scrolls each moved level's offset by its
 * origin's cell delta and plans the texels that entered it as rectangles, in
 * the reference's upload order: the whole level on first touch or on a jump of
 * res cells or more, else the new columns then the new rows, each split in two
 * where it wraps.  It writes the committed uniform to *params (levels past the
 * state's levels: origin 0, cell 1, offset 0).
 *
 * A row strip spans the whole width and a column strip the whole height.  The
 * reference writes that spanning axis from storage 0 in logical order
 * (clipmap.rs:376-377,470-471), which agrees with the sampler's
 * storage = (logical + offset) mod res (clipmap.wgsl:42-47) only while that
 * axis' offset is 0: once both axes have scrolled, its new strips hold their
 * heights rotated along their length.  That placement is the default (the
 * drop-in reproduces the reference's texture); VX_CLIP_WRAP_FIX places the
 * spanning axis at its wrapped storage position, so that every update leaves the
 * texture equal to a full rebuild at the committed origins. */
typedef struct vx_clipmap_state vx_clipmap_state;

/* Texels of one level that an update rewrites: the storage rectangle
 * (x, y, w, h), row-major, where storage texel (x+i, y+j) holds logical texel
 * ((lx0+i) mod res, (lz0+j) mod res) of a level with origin
 * (origin_x_m, origin_z_m) and cell cell_m: the height of the voxel column under
 * that texel's centre (clipmap.rs:93-170).  48 bytes. */
typedef struct vx_clip_rect {
    uint32_t level;
    uint32_t x, y, w, h;
    uint32_t lx0, lz0;
    uint32_t res;
    float    origin_x_m, origin_z_m, cell_m;
    uint32_t pad;
} vx_clip_rect;

#define VX_CLIP_WRAP_FIX     1u
#define VX_CLIP_EDGE_GUARD   16    /* clipmap.rs:207 */
#define VX_CLIP_RECTS_MAX    64u   /* 4 per level: columns and rows, each split once */

/* Clipmap::new (clipmap.rs:52-58): every level untouched. */
int  vx_clipmap_state_create(uint32_t levels, uint32_t res, float base_cell_m, float min_update_interval_s,
                             uint32_t flags, vx_clipmap_state** out);
void vx_clipmap_state_destroy(vx_clipmap_state* s);
/* Clipmap::invalidate_all (clipmap.rs:580-585): the next update rebuilds every level. */
int  vx_clipmap_invalidate(vx_clipmap_state* s);
/* Clipmap::update (clipmap.rs:182-577) without the heights: rects needs room for
 * 4 * levels entries (VX_E_INVALID otherwise, before any state changes). */
int  vx_clipmap_update(vx_clipmap_state* s, float cam_x, float cam_z, float time_s, vx_clipmap* params,
                       vx_clip_rect* rects, uint32_t cap, uint32_t* n_rects);
/* The heights of rectangles on the host (the reference's build_full_level,
 * build_col_patch and build_row_patch, clipmap.rs:107-170), rectangle after
 * rectangle, w*h floats each, row-major: ClipmapUpload::data_f32.  With
 * storage-coordinate rectangles these are vx_write_clipmap's patches. */
int  vx_clipmap_fill_rects(const vx_world_params* p, const vx_clip_rect* rects, uint32_t n, float* data);

/* ------------------------------------------------------------------------ */
/* 3. Serving                                                               */
/* ------------------------------------------------------------------------ */

/* Serving (no reference counterpart: the reference renders one window's view).
 * A server owns a device thread that runs continuous batching over the
 * context: every request queued when a launch starts joins it; all pending
 * view frames render as ONE fused batch (vx_batch_render) while all pending
 * ray batches run as ONE fused query launch on a second stream, overlapping
 * on the GPU; requests that arrive meanwhile wait for the next launch.
 *   * View streams: a client's successive frames (same stream_id) share one
 *     resident vx_view, so history (TAA, godray) carries over.  A stream's
 *     resolution is fixed by its first frame.
 *   * Admission: at submit, a new stream reserves width*height*250 bytes (its
 *     view) and a ray batch n*64 bytes (rays + hits) of memory_budget; a
 *     request that does not fit is rejected with VX_E_BUSY, and so is a frame
 *     of a stream that already has max_queued_frames frames pending.
 *   * Latency target (latency_target_ms > 0): the server fits launch time as
 *     fixed + a * view pixels + b * query rays over its recent launches
 *     (launches that create a view excluded).  Each launch takes queued
 *     requests in arrival order while its estimated time stays within a third
 *     of the target (at least one request), so one launch never holds the
 *     next for long.  A request is rejected with VX_E_BUSY when margin *
 *     (the rest of the launch in flight + the launches that the queue ahead
 *     and the request need) exceeds the target, where margin is the 99th
 *     percentile of latency / estimate over the last 256 requests admitted
 *     near the threshold (1.1 until 32 have finished): p99 latency tracks the
 *     target.  A server idle for 100 ms admits the next request regardless,
 *     so the estimate cannot lock every request out.
 *   * Edits: single-voxel edits rebuild their chunk on host worker threads
 *     (generator fill + edits + vx_build_chunk); at the next launch boundary
 *     each rebuilt chunk is written to a fresh node-arena range (best-fit),
 *     then its macro/colinfo/meta entries are rewritten and the old range is
 *     freed (double-buffered replacement).  The context's node_capacity must
 *     leave room for the replacements.
 * The world must be the one vx_host_world_build packed (slots in grid order)
 * and already uploaded to the context. */
typedef struct vx_server vx_server;
typedef struct vx_server_config {
    uint64_t memory_budget;          /* bytes for resident view streams + in-flight ray batches */
    uint32_t max_views_per_launch;   /* default 64 */
    uint32_t edit_threads;           /* host threads rebuilding edited chunks (default 8); with
                                        VX_SERVER_DEVICE_BUILD, threads filling the resident material */
    uint32_t max_queued_frames;      /* per stream, queued + in flight; more is rejected (0: no limit) */
    float    latency_target_ms;      /* submit -> done target for admission and launch size (0: none) */
    vx_world_params world;           /* generator of the resident world */
    int32_t  grid_origin[3];
    uint32_t grid_dims[3];
    uint32_t flags;                  /* VX_SERVER_* */
} vx_server_config;
/* Keep every slot's material resident in HBM (64^3 B + 16 KB per chunk, filled
 * from the generator at create on edit_threads host threads) and rebuild edited
 * chunks with the device chunk builder at launch boundaries, up to 256 chunks
 * per boundary, instead of on host worker threads. */
#define VX_SERVER_DEVICE_BUILD 1u
#define VX_REQ_VIEW 1u
#define VX_REQ_RAYS 2u
typedef struct vx_request {
    uint32_t kind;                   /* VX_REQ_VIEW or VX_REQ_RAYS */
    uint32_t stream_id;              /* VIEW: the client's view stream */
    uint32_t width, height;          /* VIEW: render = present resolution */
    vx_camera camera;                /* VIEW: this frame's camera; RAYS: the scene camera (ray
                                        batches whose scene fields differ run as separate launches) */
    const vx_ray* rays;              /* RAYS: host array, copied at submit (into the server's pinned
                                        ray pool of budget/64 rays, 64 Ki..8 Mi, when it has room) */
    uint32_t n_rays, pad;
    float*   out_rgba;               /* VIEW: optional host buffer, width*height RGBA32F final frame */
    vx_hit*  out_hits;               /* RAYS: optional host buffer of n_rays hits, written before the
                                        result is posted (ray batches post before the launch's views) */
    uint64_t tag;                    /* echoed in the result */
} vx_request;
typedef struct vx_result {
    uint64_t tag;
    uint32_t kind;
    int32_t  status;                 /* VX_OK or the launch's error */
    double   submit_ms, start_ms, done_ms;  /* server clock, ms since vx_server_create */
} vx_result;
typedef struct vx_server_stats {
    uint64_t submitted, admitted, rejected, completed, rays, edits_applied, launches, pad;
    double   view_p50_ms, view_p99_ms, rays_p50_ms, rays_p99_ms;  /* submit -> done */
    double   busy_ms, wall_ms;
    uint64_t memory_in_use, pad2;
} vx_server_stats;
/* The server counts as a live child of ctx until vx_server_destroy, which
 * stops and joins its threads (vx_destroy refuses before that). */
int  vx_server_create(vx_ctx* ctx, const vx_host_world_view* world, const vx_server_config* cfg, vx_server** out);
void vx_server_destroy(vx_server* server);
/* Queues a request (VX_E_BUSY when admission fails); never blocks on the GPU. */
int  vx_server_submit(vx_server* server, const vx_request* req);
/* Next finished request, or VX_E_AGAIN. */
int  vx_server_poll(vx_server* server, vx_result* out);
/* Single-voxel edits (world voxel coords, material). */
int  vx_server_edit(vx_server* server, const int32_t* xyz, const uint8_t* mats, uint32_t n);
/* Waits until every queued edit is on the device.  Returns early with the
 * server's error when a rebuild, upload or launch failed (errors are sticky:
 * later waits return the same code), and with VX_E_CAPACITY when the only
 * edits left are rebuilt chunks the node arena cannot place and nothing else
 * is queued that could free room. */
int  vx_server_wait_edits(vx_server* server);
/* Waits until every admitted request has finished, or returns the server's
 * sticky error as soon as one is set. */
int  vx_server_drain(vx_server* server);
/* Ends a view stream (after its queued frames) and releases its view. */
int  vx_server_close_stream(vx_server* server, uint32_t stream_id);
int  vx_server_get_stats(vx_server* server, vx_server_stats* out);

#ifdef __cplusplus
}
#endif

#endif /* VOXHIP_H */
