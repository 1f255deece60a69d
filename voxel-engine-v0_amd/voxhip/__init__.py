"""voxhip — Python bindings of libvoxhip, the MI355X implementation of the
voxel-engine-v0 primary / local TAA / godray / composite passes.

The bindings mirror the reference renderer boundary (render::Renderer,
src/render/state/mod.rs:34-712): a Context owns the resident world
(Renderer::new / apply_chunk_uploads / write_chunk_grid /
write_clipmap_updates), a View owns one camera stream and its history
(write_camera / encode_compute / resize_output).  All work runs in
libvoxhip.so; this module only moves arguments across the C ABI.  Loading
fails loudly when the library has not been built.
"""
from __future__ import annotations

import ctypes as C
import atexit
import math
import os
import weakref
from dataclasses import dataclass
from pathlib import Path

import numpy as np

LIB_PATH = Path(os.environ.get("VOXHIP_LIB") or (Path(__file__).resolve().parent / "libvoxhip.so"))

VX_OK = 0
VX_E_CAPACITY = -2
VX_E_BUSY = -6
VX_E_AGAIN = -7
REQ_VIEW, REQ_RAYS = 1, 2
INVALID = 0xFFFFFFFF
MAT_AIR, MAT_GRASS, MAT_DIRT, MAT_STONE, MAT_WOOD, MAT_LEAF, MAT_LIGHT = range(7)
MAX_CHUNK_NODES = 299593

PASS_PRIMARY = 0x01
PASS_LOCAL_TAA = 0x02
PASS_GODRAY = 0x04
PASS_COMPOSITE = 0x08
PASS_COMPOSITE_TAA = 0x10
PASS_END_FRAME = 0x80
PASS_FRAME = 0x9F
PASS_WRITE_HITS = 0x100
PASS_COUNT = 0x200
PASS_COUNT_EXECUTED = 0x400
COUNT_PARTS = ("gather", "trace", "grass", "shadow", "shade", "light", "heightfield", "textures")
NOALLOC_MAX_VIEWS = 64

OUT_COLOR, OUT_DEPTH, OUT_LOCAL, OUT_PRIMARY_HIST, OUT_SHADOW_HIST = 0, 1, 2, 3, 4
OUT_LOCAL_HIST, OUT_GODRAY, OUT_PRE_TAA, OUT_OUTPUT, OUT_HITS = 5, 6, 7, 8, 9
OUT_RAY_BYTES = 10

ARENA_NODES, ARENA_ROPES, ARENA_META, ARENA_MACRO, ARENA_COLINFO = range(5)
CTX_GODRAY_MACRO_FIX = 1
CAM_JITTER = 1
CLIP_WRAP_FIX = 1
SERVER_DEVICE_BUILD = 1
CLIP_EDGE_GUARD = 16
CLIP_RECTS_MAX = 64

NODE_DTYPE = np.dtype([("child_base", "<u4"), ("child_mask", "<u4"), ("material", "<u4"), ("key", "<u4")])
ROPES_DTYPE = np.dtype([(n, "<u4") for n in ("px", "nx", "py", "ny", "pz", "nz", "pad0", "pad1")])
META_DTYPE = np.dtype(
    [("origin", "<i4", (4,)), ("node_base", "<u4"), ("node_count", "<u4"), ("macro_base", "<u4"),
     ("colinfo_base", "<u4"), ("macro_empty", "<u4"), ("pad", "<u4", (3,))]
)
HIT_DTYPE = np.dtype(
    [("slot", "<u4"), ("node", "<u4"), ("voxel", "<i4", (3,)), ("mat", "<u4"), ("t", "<f4"), ("kind", "<u4")]
)
PROFILE_NAMES = ("voxel", "grass", "hdr", "fog", "shadow")
COUNT_NAMES = ("grid", "meta", "macro", "node", "rope", "colinfo", "clip", "tex_rd", "tex_wr")
# Bytes per counted element (gpu_types.rs struct sizes; textures count bytes).
COUNT_BYTES = (4, 48, 4, 16, 32, 4, 4, 1, 1)

assert NODE_DTYPE.itemsize == 16 and ROPES_DTYPE.itemsize == 32 and META_DTYPE.itemsize == 48
assert HIT_DTYPE.itemsize == 32


class Camera(C.Structure):
    """CameraGpu, src/render/gpu_types.rs:52-86 (464 bytes)."""

    _fields_ = [
        ("view_inv", C.c_float * 16), ("proj_inv", C.c_float * 16), ("view_proj", C.c_float * 16),
        ("prev_view_proj", C.c_float * 16), ("cam_pos", C.c_float * 4), ("ray00", C.c_float * 4),
        ("ray_dx", C.c_float * 4), ("ray_dy", C.c_float * 4), ("chunk_size", C.c_uint32),
        ("chunk_count", C.c_uint32), ("max_steps", C.c_uint32), ("frame_index", C.c_uint32),
        ("voxel_params", C.c_float * 4), ("grid_origin_chunk", C.c_int32 * 4), ("grid_dims", C.c_uint32 * 4),
        ("render_present_px", C.c_uint32 * 4), ("profile_flags", C.c_uint32), ("pad_profile", C.c_uint32 * 3),
        ("pad0", C.c_uint32 * 4), ("pad1", C.c_uint32 * 4), ("pad2", C.c_uint32 * 4),
    ]


class Clipmap(C.Structure):
    """ClipmapGpu, src/render/gpu_types.rs:105-113 (528 bytes)."""

    _fields_ = [("levels", C.c_uint32), ("res", C.c_uint32), ("base_cell_m", C.c_float), ("pad0", C.c_float),
                ("level", (C.c_float * 4) * 16), ("offset", (C.c_uint32 * 4) * 16)]


class Caps(C.Structure):
    _fields_ = [("node_capacity", C.c_uint64), ("chunk_capacity", C.c_uint32), ("grid_capacity", C.c_uint32),
                ("clip_levels", C.c_uint32), ("clip_res", C.c_uint32), ("device", C.c_int32), ("flags", C.c_uint32)]


class Region(C.Structure):
    _fields_ = [("offset", C.c_uint64), ("count", C.c_uint64), ("data", C.c_void_p)]


class ClipPatch(C.Structure):
    _fields_ = [("level", C.c_uint32), ("x", C.c_uint32), ("y", C.c_uint32), ("w", C.c_uint32), ("h", C.c_uint32),
                ("data", C.c_void_p)]


class ClipRect(C.Structure):
    """vx_clip_rect: storage texels (x, y, w, h) of one level holding logical
    texels ((lx0 + i) mod res, (lz0 + j) mod res) of a level with the given
    origin and cell (48 bytes)."""

    _fields_ = [("level", C.c_uint32), ("x", C.c_uint32), ("y", C.c_uint32), ("w", C.c_uint32), ("h", C.c_uint32),
                ("lx0", C.c_uint32), ("lz0", C.c_uint32), ("res", C.c_uint32), ("origin_x_m", C.c_float),
                ("origin_z_m", C.c_float), ("cell_m", C.c_float), ("pad", C.c_uint32)]

    def as_tuple(self) -> tuple:
        return (self.level, self.x, self.y, self.w, self.h, self.lx0, self.lz0, self.res, self.origin_x_m,
                self.origin_z_m, self.cell_m)


class WorldParams(C.Structure):
    _fields_ = [("seed", C.c_uint32), ("base_y", C.c_int32), ("amplitude", C.c_int32), ("dirt_depth", C.c_int32),
                ("tree_density", C.c_float), ("trunk_min", C.c_int32), ("trunk_max", C.c_int32),
                ("leaf_radius", C.c_int32), ("wavelength0", C.c_int32)]


class Pose(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("yaw", C.c_float), ("pitch", C.c_float), ("fovy", C.c_float),
                ("z_near", C.c_float), ("z_far", C.c_float)]


class BatchConfig(C.Structure):
    _fields_ = [("waves_per_simd", C.c_uint32), ("grab", C.c_uint32), ("use_graph", C.c_uint32),
                ("tile_shape", C.c_uint32)]

    def as_dict(self) -> dict:
        return {"waves_per_simd": self.waves_per_simd, "grab": self.grab, "use_graph": self.use_graph,
                "tile_shape": TILE_SHAPES[self.tile_shape] if self.tile_shape < len(TILE_SHAPES) else self.tile_shape}


TILE_SHAPES = ("8x8", "16x4", "32x2")  # vx_batch_config.tile_shape


class QueryConfig(C.Structure):
    _fields_ = [("waves_per_simd", C.c_uint32), ("refill_below", C.c_uint32), ("sort_rays", C.c_uint32),
                ("pad", C.c_uint32)]

    def as_dict(self) -> dict:
        return {"waves_per_simd": self.waves_per_simd, "refill_below": self.refill_below, "sort_rays": self.sort_rays}


class ServerConfig(C.Structure):
    _fields_ = [("memory_budget", C.c_uint64), ("max_views_per_launch", C.c_uint32), ("edit_threads", C.c_uint32),
                ("max_queued_frames", C.c_uint32), ("latency_target_ms", C.c_float), ("world", WorldParams),
                ("grid_origin", C.c_int32 * 3), ("grid_dims", C.c_uint32 * 3), ("flags", C.c_uint32)]


class Request(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("stream_id", C.c_uint32), ("width", C.c_uint32), ("height", C.c_uint32),
                ("camera", Camera), ("rays", C.c_void_p), ("n_rays", C.c_uint32), ("pad", C.c_uint32),
                ("out_rgba", C.c_void_p), ("out_hits", C.c_void_p), ("tag", C.c_uint64)]


class Result(C.Structure):
    _fields_ = [("tag", C.c_uint64), ("kind", C.c_uint32), ("status", C.c_int32), ("submit_ms", C.c_double),
                ("start_ms", C.c_double), ("done_ms", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("submitted", "admitted", "rejected", "completed", "rays", "edits_applied",
                                          "launches", "pad")] + \
               [(n, C.c_double) for n in ("view_p50_ms", "view_p99_ms", "rays_p50_ms", "rays_p99_ms", "busy_ms",
                                          "wall_ms")] + [("memory_in_use", C.c_uint64), ("pad2", C.c_uint64)]

    def as_dict(self) -> dict:
        return {n: getattr(self, n) for n, _ in self._fields_ if not n.startswith("pad")}


class HostWorldView(C.Structure):
    _fields_ = [("nodes", C.c_void_p), ("node_count", C.c_uint64), ("ropes", C.c_void_p), ("meta", C.c_void_p),
                ("chunk_count", C.c_uint32), ("macro_words", C.c_void_p), ("colinfo_words", C.c_void_p),
                ("grid", C.c_void_p), ("grid_cells", C.c_uint32), ("grid_origin", C.c_int32 * 3),
                ("grid_dims", C.c_uint32 * 3)]


assert C.sizeof(Camera) == 464 and C.sizeof(Clipmap) == 528 and C.sizeof(ClipRect) == 48


class VoxhipError(RuntimeError):
    pass


def _load() -> C.CDLL:
    if not LIB_PATH.exists():
        raise VoxhipError(f"{LIB_PATH} is missing: build it with `python voxel-engine-v0_amd/build.py`")
    lib = C.CDLL(str(LIB_PATH))
    vp, u32, i32, u64, f32 = C.c_void_p, C.c_uint32, C.c_int32, C.c_uint64, C.c_float
    sig = {
        "vx_last_error": ([], C.c_char_p), "vx_version": ([], C.c_char_p), "vx_device_count": ([], C.c_int),
        "vx_create": ([C.POINTER(Caps), C.POINTER(vp)], C.c_int), "vx_destroy": ([vp], C.c_int),
        "vx_upload_regions": ([vp, C.c_int, C.POINTER(Region), u32], C.c_int),
        "vx_write_grid": ([vp, vp, u32], C.c_int),
        "vx_write_clipmap": ([vp, C.POINTER(Clipmap), C.POINTER(ClipPatch), u32], C.c_int),
        "vx_view_create": ([vp, u32, u32, u32, u32, C.POINTER(vp)], C.c_int), "vx_view_destroy": ([vp], None),
        "vx_view_resize": ([vp, u32, u32, u32, u32], C.c_int), "vx_view_reset": ([vp], C.c_int),
        "vx_write_camera": ([vp, C.POINTER(Camera)], C.c_int), "vx_render": ([vp, u32, vp], C.c_int),
        "vx_read_output": ([vp, C.c_int, vp, C.c_size_t], C.c_int),
        "vx_output_device_ptr": ([vp, C.c_int, C.POINTER(vp), C.POINTER(C.c_size_t)], C.c_int),
        "vx_write_history": ([vp, C.c_int, vp, C.c_size_t], C.c_int),
        "vx_view_counters": ([vp, C.POINTER(u64), C.c_int], C.c_int),
        "vx_view_timings": ([vp, C.POINTER(f32)], C.c_int),
        "vx_view_profile": ([vp, C.POINTER(u32), C.c_int], C.c_int),
        "vx_view_count_parts": ([vp, C.POINTER(u64)], C.c_int),
        "vx_render_views_noalloc": ([vp, C.POINTER(vp), u32, u32, vp], C.c_int),
        "vx_query_rays": ([vp, C.POINTER(Camera), vp, u32, vp, vp], C.c_int),
        "vx_query_timing": ([vp, C.POINTER(f32)], C.c_int),
        "vx_query_count": ([vp, C.POINTER(Camera), vp, u32, vp, C.c_int, C.POINTER(u64), C.POINTER(u64)], C.c_int),
        "vx_query_set_config": ([vp, C.POINTER(QueryConfig)], C.c_int),
        "vx_query_get_config": ([vp, C.POINTER(QueryConfig)], C.c_int),
        "vx_query_tune": ([vp, C.POINTER(Camera), vp, u32, vp, C.POINTER(QueryConfig)], C.c_int),
        "vx_sync": ([vp], C.c_int),
        "vx_server_create": ([vp, C.POINTER(HostWorldView), C.POINTER(ServerConfig), C.POINTER(vp)], C.c_int),
        "vx_server_destroy": ([vp], None), "vx_server_submit": ([vp, C.POINTER(Request)], C.c_int),
        "vx_server_poll": ([vp, C.POINTER(Result)], C.c_int),
        "vx_server_edit": ([vp, vp, vp, u32], C.c_int), "vx_server_wait_edits": ([vp], C.c_int),
        "vx_server_drain": ([vp], C.c_int), "vx_server_close_stream": ([vp, u32], C.c_int),
        "vx_server_get_stats": ([vp, C.POINTER(ServerStats)], C.c_int),
        "vx_batch_create": ([vp, C.POINTER(vp), u32, C.POINTER(vp)], C.c_int), "vx_batch_destroy": ([vp], None),
        "vx_batch_render": ([vp, u32, vp], C.c_int), "vx_batch_timing": ([vp, C.POINTER(f32)], C.c_int),
        "vx_batch_set_config": ([vp, C.POINTER(BatchConfig)], C.c_int),
        "vx_batch_get_config": ([vp, C.POINTER(BatchConfig)], C.c_int),
        "vx_batch_tune": ([vp, u32, u32, C.POINTER(BatchConfig)], C.c_int),
        "vx_batch_set_rays": ([vp, vp, vp, u32, vp], C.c_int),
        "vx_build_chunk": ([vp, vp, i32, i32, vp, vp, u32, C.POINTER(u32), vp, vp], C.c_int),
        "vx_world_default": ([C.POINTER(WorldParams)], None),
        "vx_world_ground_height": ([C.POINTER(WorldParams), i32, i32], i32),
        "vx_world_fill_chunk": ([C.POINTER(WorldParams), i32, i32, i32, vp, vp], C.c_int),
        "vx_host_world_build": ([C.POINTER(WorldParams), C.POINTER(i32), C.POINTER(u32), u32, C.POINTER(vp)], C.c_int),
        "vx_host_world_free": ([vp], None),
        "vx_host_world_get": ([vp, C.POINTER(HostWorldView)], C.c_int),
        "vx_host_world_edit": ([vp, vp, vp, u32, vp, C.POINTER(u32)], C.c_int),
        "vx_camera_build": ([C.POINTER(Pose), u32, u32, u32, u32, u32, f32, f32, f32, C.POINTER(i32), C.POINTER(u32), u32,
                             vp, u32, C.POINTER(Camera)], C.c_int),
        "vx_clipmap_build": ([C.POINTER(WorldParams), f32, f32, u32, u32, f32, C.POINTER(Clipmap), vp], C.c_int),
        "vx_clipmap_state_create": ([u32, u32, f32, f32, u32, C.POINTER(vp)], C.c_int),
        "vx_clipmap_state_destroy": ([vp], None), "vx_clipmap_invalidate": ([vp], C.c_int),
        "vx_clipmap_update": ([vp, f32, f32, f32, C.POINTER(Clipmap), C.POINTER(ClipRect), u32, C.POINTER(u32)],
                              C.c_int),
        "vx_clipmap_fill_rects": ([C.POINTER(WorldParams), C.POINTER(ClipRect), u32, vp], C.c_int),
        "vx_write_clipmap_rects": ([vp, C.POINTER(Clipmap), C.POINTER(WorldParams), C.POINTER(ClipRect), u32], C.c_int),
        "vx_read_clipmap": ([vp, vp, C.c_size_t], C.c_int),
        "vx_read_region": ([vp, C.c_int, u64, u64, vp], C.c_int),
        "vx_chunk_builder_create": ([vp, u32, C.POINTER(vp)], C.c_int),
        "vx_chunk_builder_destroy": ([vp], None),
        "vx_chunk_builder_count": ([vp, u32, vp, vp, C.POINTER(i32), i32, C.POINTER(u32)], C.c_int),
        "vx_chunk_builder_emit": ([vp, C.POINTER(u64), C.POINTER(u32), C.POINTER(u32)], C.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    return lib


LIB = _load()
EXPORTED = (
    "vx_last_error", "vx_version", "vx_device_count", "vx_create", "vx_destroy", "vx_upload_regions", "vx_write_grid",
    "vx_write_clipmap", "vx_view_create", "vx_view_destroy", "vx_view_resize", "vx_view_reset", "vx_write_camera",
    "vx_render", "vx_read_output", "vx_output_device_ptr", "vx_write_history", "vx_view_counters", "vx_view_timings", "vx_view_profile", "vx_render_views_noalloc", "vx_view_count_parts",
    "vx_server_create", "vx_server_destroy", "vx_server_submit", "vx_server_poll", "vx_server_edit",
    "vx_server_wait_edits", "vx_server_drain", "vx_server_close_stream", "vx_server_get_stats",
    "vx_query_rays", "vx_query_timing", "vx_query_count", "vx_query_set_config", "vx_query_get_config", "vx_query_tune", "vx_sync", "vx_batch_create", "vx_batch_destroy", "vx_batch_render", "vx_batch_timing",
    "vx_batch_set_config", "vx_batch_get_config", "vx_batch_tune", "vx_batch_set_rays", "vx_build_chunk", "vx_world_default", "vx_world_ground_height", "vx_world_fill_chunk",
    "vx_host_world_build", "vx_host_world_free", "vx_host_world_get", "vx_host_world_edit", "vx_camera_build",
    "vx_clipmap_build", "vx_clipmap_state_create", "vx_clipmap_state_destroy", "vx_clipmap_invalidate",
    "vx_clipmap_update", "vx_clipmap_fill_rects", "vx_write_clipmap_rects", "vx_read_clipmap", "vx_read_region",
    "vx_chunk_builder_create", "vx_chunk_builder_destroy", "vx_chunk_builder_count", "vx_chunk_builder_emit",
)


def _check(rc: int, what: str) -> None:
    if rc != VX_OK:
        raise VoxhipError(f"{what} failed ({rc}): {LIB.vx_last_error().decode()}")


def _ptr(a: np.ndarray) -> int:
    assert a.flags.c_contiguous
    return a.ctypes.data


def device_count() -> int:
    return int(LIB.vx_device_count())


# ---------------------------------------------------------------------------
# Host producers
# ---------------------------------------------------------------------------
def world_params(seed: int = 42, **kw) -> WorldParams:
    p = WorldParams()
    LIB.vx_world_default(C.byref(p))
    p.seed = seed
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def ground_height(p: WorldParams, wx: int, wz: int) -> int:
    return int(LIB.vx_world_ground_height(C.byref(p), wx, wz))


def fill_chunk(p: WorldParams, cx: int, cy: int, cz: int) -> tuple[np.ndarray, np.ndarray]:
    mat = np.zeros(64 * 64 * 64, np.uint8)
    ground = np.zeros(64 * 64, np.int32)
    _check(LIB.vx_world_fill_chunk(C.byref(p), cx, cy, cz, _ptr(mat), _ptr(ground)), "vx_world_fill_chunk")
    return mat, ground


@dataclass
class ChunkBuild:
    nodes: np.ndarray
    ropes: np.ndarray
    macro: np.ndarray
    colinfo: np.ndarray


def build_chunk(material: np.ndarray, ground: np.ndarray, chunk_oy: int, dirt_depth: int,
                node_capacity: int = MAX_CHUNK_NODES) -> ChunkBuild:
    material = np.ascontiguousarray(material, np.uint8).reshape(-1)
    ground = np.ascontiguousarray(ground, np.int32).reshape(-1)
    assert material.size == 64 ** 3 and ground.size == 64 * 64
    nodes = np.zeros(node_capacity, NODE_DTYPE)
    ropes = np.zeros(node_capacity, ROPES_DTYPE)
    macro = np.zeros(16, np.uint32)
    colinfo = np.zeros(2048, np.uint32)
    n = C.c_uint32(0)
    _check(LIB.vx_build_chunk(_ptr(material), _ptr(ground), chunk_oy, dirt_depth, _ptr(nodes), _ptr(ropes),
                              node_capacity, C.byref(n), _ptr(macro), _ptr(colinfo)), "vx_build_chunk")
    return ChunkBuild(nodes[: n.value].copy(), ropes[: n.value].copy(), macro, colinfo)


class HostWorld:
    """A resident world built on the host (every grid chunk through vx_build_chunk)."""

    def __init__(self, params: WorldParams, grid_origin=(0, 0, 0), grid_dims=(8, 4, 8), threads: int = 0):
        self.params = params
        h = C.c_void_p()
        go = (C.c_int32 * 3)(*grid_origin)
        gd = (C.c_uint32 * 3)(*grid_dims)
        _check(LIB.vx_host_world_build(C.byref(params), go, gd, threads, C.byref(h)), "vx_host_world_build")
        self._h = h
        self._refresh()

    def _refresh(self) -> None:
        v = HostWorldView()
        _check(LIB.vx_host_world_get(self._h, C.byref(v)), "vx_host_world_get")

        def arr(ptr, count, dtype):
            if count == 0:
                return np.zeros(0, dtype)
            buf = (C.c_byte * (count * np.dtype(dtype).itemsize)).from_address(ptr)
            return np.frombuffer(buf, dtype=dtype, count=count)

        self.nodes = arr(v.nodes, v.node_count, NODE_DTYPE)
        self.ropes = arr(v.ropes, v.node_count, ROPES_DTYPE)
        self.meta = arr(v.meta, v.chunk_count, META_DTYPE)
        self.macro = arr(v.macro_words, v.chunk_count * 16, np.uint32)
        self.colinfo = arr(v.colinfo_words, v.chunk_count * 2048, np.uint32)
        self.grid = arr(v.grid, v.grid_cells, np.uint32)
        self.grid_origin = tuple(v.grid_origin)
        self.grid_dims = tuple(v.grid_dims)
        self.chunk_count = int(v.chunk_count)

    def edit(self, xyz: np.ndarray, mats: np.ndarray) -> list[int]:
        xyz = np.ascontiguousarray(xyz, np.int32).reshape(-1, 3)
        mats = np.ascontiguousarray(mats, np.uint8).reshape(-1)
        out = np.zeros(max(1, len(mats)), np.uint32)
        n = C.c_uint32(0)
        _check(LIB.vx_host_world_edit(self._h, _ptr(xyz), _ptr(mats), len(mats), _ptr(out), C.byref(n)),
               "vx_host_world_edit")
        self._refresh()
        return [int(s) for s in out[: n.value]]

    def close(self) -> None:
        if self._h:
            LIB.vx_host_world_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def camera_build(pose: Pose, render_wh, present_wh, frame_index: int, *, time_s: float = 0.0,
                 wind_strength: float = 2.0, fog_density: float = 0.002, grid_origin=(0, 0, 0),
                 grid_dims=(8, 4, 8), chunk_count: int = 0, prev_view_proj=None, jitter: bool = False) -> Camera:
    cam = Camera()
    go = (C.c_int32 * 3)(*grid_origin)
    gd = (C.c_uint32 * 3)(*grid_dims)
    prev = None
    if prev_view_proj is not None:
        pv = np.ascontiguousarray(prev_view_proj, np.float32).reshape(16)
        prev = pv.ctypes.data
    _check(LIB.vx_camera_build(C.byref(pose), render_wh[0], render_wh[1], present_wh[0], present_wh[1], frame_index,
                               time_s, wind_strength, fog_density, go, gd, chunk_count, prev,
                               CAM_JITTER if jitter else 0, C.byref(cam)), "vx_camera_build")
    return cam


def make_pose(pos, yaw: float, pitch: float, fovy_deg: float = 70.0, z_near: float = 0.1, z_far: float = 1000.0) -> Pose:
    p = Pose()
    p.pos[0], p.pos[1], p.pos[2] = (float(v) for v in pos)
    p.yaw, p.pitch, p.fovy = yaw, pitch, math.radians(fovy_deg)
    p.z_near, p.z_far = z_near, z_far
    return p


def clipmap_build(params: WorldParams, cam_x: float, cam_z: float, levels: int = 5, res: int = 256,
                  base_cell_m: float = 1.0) -> tuple[Clipmap, np.ndarray]:
    clip = Clipmap()
    heights = np.zeros((max(levels, 1), res, res), np.float32)
    _check(LIB.vx_clipmap_build(C.byref(params), cam_x, cam_z, levels, res, base_cell_m, C.byref(clip), _ptr(heights)),
           "vx_clipmap_build")
    return clip, heights


class ClipmapState:
    """Clipmap (src/clipmap.rs:44-60) with its per-frame update
    (clipmap.rs:182-577): committed origin cell and toroidal offset per level.
    wrap_fix places a strip's spanning axis at its wrapped storage position
    (VX_CLIP_WRAP_FIX); by default strips are placed as the reference does."""

    def __init__(self, levels: int = 5, res: int = 256, base_cell_m: float = 1.0, min_update_interval_s: float = 0.0,
                 wrap_fix: bool = False):
        h = C.c_void_p()
        _check(LIB.vx_clipmap_state_create(levels, res, base_cell_m, min_update_interval_s,
                                           CLIP_WRAP_FIX if wrap_fix else 0, C.byref(h)), "vx_clipmap_state_create")
        self._h = h
        self.levels, self.res = levels, res
        self._rects = (ClipRect * (4 * levels))()

    def update(self, cam_x: float, cam_z: float, time_s: float) -> tuple[Clipmap, list]:
        """Clipmap::update without the heights: (committed uniform, rectangles)."""
        clip = Clipmap()
        n = C.c_uint32()
        _check(LIB.vx_clipmap_update(self._h, cam_x, cam_z, time_s, C.byref(clip), self._rects, len(self._rects),
                                     C.byref(n)), "vx_clipmap_update")
        return clip, [ClipRect.from_buffer_copy(self._rects[i]) for i in range(n.value)]

    def invalidate(self) -> None:
        _check(LIB.vx_clipmap_invalidate(self._h), "vx_clipmap_invalidate")

    def close(self) -> None:
        if self._h:
            LIB.vx_clipmap_state_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _rect_array(rects: list):
    arr = (ClipRect * max(1, len(rects)))()
    for i, r in enumerate(rects):
        arr[i] = r
    return arr


def clipmap_fill_rects(params: WorldParams, rects: list) -> list:
    """Heights of each rectangle on the host (ClipmapUpload::data_f32): a list
    of (h, w) float32 arrays."""
    total = sum(r.w * r.h for r in rects)
    data = np.zeros(max(total, 1), np.float32)
    _check(LIB.vx_clipmap_fill_rects(C.byref(params), _rect_array(rects), len(rects), _ptr(data)),
           "vx_clipmap_fill_rects")
    out, k = [], 0
    for r in rects:
        out.append(data[k:k + r.w * r.h].reshape(r.h, r.w))
        k += r.w * r.h
    return out


def apply_clip_rects(heights: np.ndarray, rects: list, data: list) -> None:
    """Writes filled rectangles into a (levels, res, res) storage array."""
    for r, d in zip(rects, data):
        heights[r.level, r.y:r.y + r.h, r.x:r.x + r.w] = d


def clip_logical(clip: Clipmap, heights: np.ndarray, level: int) -> np.ndarray:
    """A level of a storage array in logical order: logical (lx, lz) is
    storage ((lx + off_x) mod res, (lz + off_z) mod res)."""
    res = clip.res
    ox, oz = clip.offset[level][0], clip.offset[level][1]
    return np.roll(np.roll(heights[level], -int(oz), axis=0), -int(ox), axis=1)[:res, :res]


# ---------------------------------------------------------------------------
# Renderer boundary
# ---------------------------------------------------------------------------
class Context:
    """Resident world on one GPU (Renderer::new + uploads)."""

    def __init__(self, node_capacity: int, chunk_capacity: int, grid_capacity: int, clip_levels: int = 5,
                 clip_res: int = 256, device: int = 0, godray_macro_fix: bool = False):
        flags = CTX_GODRAY_MACRO_FIX if godray_macro_fix else 0
        caps = Caps(node_capacity, chunk_capacity, grid_capacity, clip_levels, clip_res, device, flags)
        h = C.c_void_p()
        _check(LIB.vx_create(C.byref(caps), C.byref(h)), "vx_create")
        self._h = h
        self.clip_levels, self.clip_res = clip_levels, clip_res

    def upload(self, arena: int, array: np.ndarray, offset: int = 0) -> None:
        a = np.ascontiguousarray(array)
        r = Region(offset, len(a), a.ctypes.data)
        _check(LIB.vx_upload_regions(self._h, arena, C.byref(r), 1), "vx_upload_regions")

    def upload_world(self, w: HostWorld) -> None:
        self.upload(ARENA_NODES, w.nodes)
        self.upload(ARENA_ROPES, w.ropes)
        self.upload(ARENA_META, w.meta)
        self.upload(ARENA_MACRO, w.macro)
        self.upload(ARENA_COLINFO, w.colinfo)
        self.write_grid(w.grid)

    def write_grid(self, grid: np.ndarray) -> None:
        g = np.ascontiguousarray(grid, np.uint32)
        _check(LIB.vx_write_grid(self._h, g.ctypes.data, len(g)), "vx_write_grid")

    def write_clipmap(self, clip: Clipmap, heights: np.ndarray) -> None:
        h = np.ascontiguousarray(heights, np.float32)
        patches = (ClipPatch * max(1, clip.levels))()
        for lv in range(clip.levels):
            patches[lv] = ClipPatch(lv, 0, 0, clip.res, clip.res, h[lv].ctypes.data)
        _check(LIB.vx_write_clipmap(self._h, C.byref(clip), patches, clip.levels), "vx_write_clipmap")

    def write_clipmap_patches(self, clip: Clipmap, rects: list, data: list) -> None:
        """Renderer::write_clipmap_updates with host-built rectangles."""
        keep = [np.ascontiguousarray(d, np.float32) for d in data]
        patches = (ClipPatch * max(1, len(rects)))()
        for i, (r, d) in enumerate(zip(rects, keep)):
            patches[i] = ClipPatch(r.level, r.x, r.y, r.w, r.h, d.ctypes.data)
        _check(LIB.vx_write_clipmap(self._h, C.byref(clip), patches, len(rects)), "vx_write_clipmap")

    def write_clipmap_rects(self, clip: Clipmap, params: WorldParams, rects: list) -> None:
        """Renderer::write_clipmap_updates with the heights generated on the device."""
        _check(LIB.vx_write_clipmap_rects(self._h, C.byref(clip), C.byref(params), _rect_array(rects), len(rects)),
               "vx_write_clipmap_rects")

    def read_region(self, arena: int, offset: int, count: int) -> np.ndarray:
        """Elements [offset, offset + count) of an arena, copied to the host."""
        dt = {ARENA_NODES: NODE_DTYPE, ARENA_ROPES: ROPES_DTYPE, ARENA_META: META_DTYPE, ARENA_MACRO: np.uint32,
              ARENA_COLINFO: np.uint32}[arena]
        out = np.zeros(count, dt)
        _check(LIB.vx_read_region(self._h, arena, offset, count, out.ctypes.data), "vx_read_region")
        return out

    def read_clipmap(self) -> np.ndarray:
        out = np.zeros((self.clip_levels, self.clip_res, self.clip_res), np.float32)
        _check(LIB.vx_read_clipmap(self._h, _ptr(out), out.nbytes), "vx_read_clipmap")
        return out

    def sync(self) -> None:
        _check(LIB.vx_sync(self._h), "vx_sync")

    def close(self) -> None:
        """Frees the resident world.  The library refuses (VX_E_BUSY) while a
        view, batch, server or chunk builder of this context is alive; the
        context then stays open and the error is raised."""
        if self._h:
            _check(LIB.vx_destroy(self._h), "vx_destroy")
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class View:
    """One camera stream with its own history textures."""

    def __init__(self, ctx: Context, present_wh, render_wh=None):
        render_wh = render_wh or present_wh
        h = C.c_void_p()
        _check(LIB.vx_view_create(ctx._h, present_wh[0], present_wh[1], render_wh[0], render_wh[1], C.byref(h)),
               "vx_view_create")
        self._h = h
        self.ctx = ctx
        self.pw, self.ph = present_wh
        self.rw, self.rh = render_wh

    def write_camera(self, cam: Camera) -> None:
        _check(LIB.vx_write_camera(self._h, C.byref(cam)), "vx_write_camera")

    def render(self, mask: int = PASS_FRAME, stream: int | None = None) -> None:
        _check(LIB.vx_render(self._h, mask, stream), "vx_render")

    def reset(self) -> None:
        _check(LIB.vx_view_reset(self._h), "vx_view_reset")

    def read(self, which: int) -> np.ndarray:
        r, p = self.rw * self.rh, self.pw * self.ph
        shapes = {
            OUT_COLOR: ((self.rh, self.rw, 4), np.float32), OUT_DEPTH: ((self.rh, self.rw), np.float32),
            OUT_LOCAL: ((self.rh, self.rw, 4), np.float32), OUT_PRIMARY_HIST: ((self.rh, self.rw, 4), np.float32),
            OUT_SHADOW_HIST: ((self.rh, self.rw), np.float32), OUT_LOCAL_HIST: ((self.rh, self.rw, 4), np.float32),
            OUT_GODRAY: ((self.rh, self.rw, 4), np.float32), OUT_PRE_TAA: ((self.ph, self.pw, 4), np.float32),
            OUT_OUTPUT: ((self.ph, self.pw, 4), np.float32), OUT_HITS: ((self.rh, self.rw), HIT_DTYPE),
            OUT_RAY_BYTES: ((self.rh, self.rw), np.float32),
        }
        shape, dt = shapes[which]
        out = np.empty(shape, dt)
        _check(LIB.vx_read_output(self._h, which, out.ctypes.data, out.nbytes), "vx_read_output")
        del r, p
        return out

    def write_history(self, which: int, data: np.ndarray) -> None:
        d = np.ascontiguousarray(data)
        _check(LIB.vx_write_history(self._h, which, d.ctypes.data, d.nbytes), "vx_write_history")

    def counters(self, reset: bool = True) -> dict:
        out = (C.c_uint64 * 9)()
        _check(LIB.vx_view_counters(self._h, out, 1 if reset else 0), "vx_view_counters")
        return dict(zip(COUNT_NAMES, (int(v) for v in out)))

    def count_parts(self) -> dict:
        """Counted bytes by part of the primary pass (vx_view_count_parts); cleared by counters(reset=True)."""
        out = (C.c_uint64 * len(COUNT_PARTS))()
        _check(LIB.vx_view_count_parts(self._h, out), "vx_view_count_parts")
        return dict(zip(COUNT_PARTS, (int(v) for v in out)))

    def profile(self, reset: bool = True) -> dict:
        """The primary pass's profile counters (vx_view_profile)."""
        out = (C.c_uint32 * 5)()
        _check(LIB.vx_view_profile(self._h, out, 1 if reset else 0), "vx_view_profile")
        return dict(zip(PROFILE_NAMES, (int(v) for v in out)))

    def timings(self) -> list[float]:
        ms = (C.c_float * 5)()
        _check(LIB.vx_view_timings(self._h, ms), "vx_view_timings")
        return [float(v) for v in ms]

    def close(self) -> None:
        if self._h:
            LIB.vx_view_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class HipStream:
    """A non-blocking HIP stream at the default priority (for passing to the C
    ABI calls that take a stream)."""

    def __init__(self):
        rt = DeviceArray._rt()
        rt.hipStreamCreateWithFlags.argtypes = [C.POINTER(C.c_void_p), C.c_uint]
        rt.hipStreamSynchronize.argtypes = [C.c_void_p]
        rt.hipStreamDestroy.argtypes = [C.c_void_p]
        self.handle = C.c_void_p()
        if rt.hipStreamCreateWithFlags(C.byref(self.handle), 1) != 0:  # hipStreamNonBlocking
            raise VoxhipError("hipStreamCreateWithFlags failed")

    def sync(self) -> None:
        if DeviceArray._rt().hipStreamSynchronize(self.handle) != 0:
            raise VoxhipError("hipStreamSynchronize failed")

    def close(self) -> None:
        if self.handle:
            DeviceArray._rt().hipStreamDestroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ChunkBuilder:
    """The chunk builder on the device (vx_chunk_builder_*): count a batch of
    chunks whose material (n x 64^3 u8, y-slab) and ground (n x 64^2 i32) are
    device arrays, then emit them into the context's arenas at the given bases."""

    def __init__(self, ctx: Context, max_chunks: int):
        h = C.c_void_p()
        _check(LIB.vx_chunk_builder_create(ctx._h, max_chunks, C.byref(h)), "vx_chunk_builder_create")
        self._h = h
        self.ctx = ctx
        self.max_chunks = max_chunks

    def count(self, material: "DeviceArray", ground: "DeviceArray", chunk_oy, dirt_depth: int,
              n: int | None = None) -> np.ndarray:
        oy = np.ascontiguousarray(chunk_oy, np.int32)
        n = len(oy) if n is None else n
        counts = np.zeros(max(n, 1), np.uint32)
        _check(LIB.vx_chunk_builder_count(self._h, n, material.ptr, ground.ptr,
                                          oy.ctypes.data_as(C.POINTER(C.c_int32)), dirt_depth,
                                          counts.ctypes.data_as(C.POINTER(C.c_uint32))), "vx_chunk_builder_count")
        return counts[:n]

    def emit(self, node_base, macro_base, colinfo_base) -> None:
        nb = np.ascontiguousarray(node_base, np.uint64)
        mb = np.ascontiguousarray(macro_base, np.uint32)
        cb = np.ascontiguousarray(colinfo_base, np.uint32)
        _check(LIB.vx_chunk_builder_emit(self._h, nb.ctypes.data_as(C.POINTER(C.c_uint64)),
                                         mb.ctypes.data_as(C.POINTER(C.c_uint32)),
                                         cb.ctypes.data_as(C.POINTER(C.c_uint32))), "vx_chunk_builder_emit")

    def close(self) -> None:
        if self._h:
            LIB.vx_chunk_builder_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceArray:
    """A device allocation through the HIP runtime (for the C ABI calls that take
    device pointers, e.g. ray queries); upload / download copy synchronously."""

    _hip = None

    @classmethod
    def _rt(cls):
        if cls._hip is None:
            cls._hip = C.CDLL("libamdhip64.so")
            cls._hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            cls._hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            cls._hip.hipFree.argtypes = [C.c_void_p]
        return cls._hip

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        self.ptr = C.c_void_p()
        self.owner = True
        if self._rt().hipMalloc(C.byref(self.ptr), max(self.nbytes, 1)) != 0:
            raise VoxhipError(f"hipMalloc({self.nbytes}) failed")

    def view(self, offset: int, nbytes: int) -> "DeviceArray":
        """A non-owning sub-range (the parent must outlive it)."""
        assert 0 <= offset and offset + nbytes <= self.nbytes
        v = DeviceArray.__new__(DeviceArray)
        v.nbytes, v.owner, v.ptr = int(nbytes), False, C.c_void_p(self.ptr.value + int(offset))
        return v

    @classmethod
    def from_numpy(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes)
        d.upload(a)
        return d

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        if self._rt().hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, 1) != 0:  # host to device
            raise VoxhipError("hipMemcpy to device failed")

    def download(self, out: np.ndarray) -> np.ndarray:
        assert out.flags.c_contiguous and out.nbytes <= self.nbytes
        if self._rt().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2) != 0:  # device to host
            raise VoxhipError("hipMemcpy to host failed")
        return out

    def close(self) -> None:
        if self.ptr and self.owner:
            self._rt().hipFree(self.ptr)
        self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def query_rays_device(ctx: "Context", scene_cam: Camera, rays: DeviceArray, hits: DeviceArray, n: int,
                      stream: int | None = None) -> None:
    """vx_query_rays on device arrays (stream-ordered; does not wait)."""
    _check(LIB.vx_query_rays(ctx._h, C.byref(scene_cam), rays.ptr, n, hits.ptr, stream), "vx_query_rays")


def query_rays(ctx: "Context", scene_cam: Camera, rays: np.ndarray) -> np.ndarray:
    """First voxel hit of each ray ((n, 8) float32 as vx_ray); returns vx_hit records."""
    rays = np.ascontiguousarray(rays, np.float32)
    n = len(rays)
    d_rays = DeviceArray.from_numpy(rays)
    d_hits = DeviceArray(n * HIT_DTYPE.itemsize)
    query_rays_device(ctx, scene_cam, d_rays, d_hits, n)
    ctx.sync()
    out = d_hits.download(np.empty(n, HIT_DTYPE))
    d_rays.close()
    d_hits.close()
    return out


def query_count(ctx: "Context", scene_cam: Camera, rays: DeviceArray, hits: DeviceArray, n: int,
                executed: bool = False) -> tuple[dict, dict]:
    """vx_query_count: the load counts of a query over n device rays (the
    reference algorithm's, or with executed those vx_query_rays makes), as
    (counts by COUNT_NAMES, bytes by COUNT_PARTS).  Synchronous."""
    out = (C.c_uint64 * len(COUNT_NAMES))()
    parts = (C.c_uint64 * len(COUNT_PARTS))()
    _check(LIB.vx_query_count(ctx._h, C.byref(scene_cam), rays.ptr, n, hits.ptr, 1 if executed else 0, out, parts),
           "vx_query_count")
    return dict(zip(COUNT_NAMES, (int(v) for v in out))), dict(zip(COUNT_PARTS, (int(v) for v in parts)))


def query_timing(ctx: "Context") -> float:
    ms = C.c_float()
    _check(LIB.vx_query_timing(ctx._h, C.byref(ms)), "vx_query_timing")
    return float(ms.value)


def query_config(ctx: "Context", waves_per_simd: int | None = None, refill_below: int | None = None,
                 sort_rays: bool | None = None) -> QueryConfig:
    """Reads (and, given values, sets) the context's ray-query configuration."""
    cfg = QueryConfig()
    _check(LIB.vx_query_get_config(ctx._h, C.byref(cfg)), "vx_query_get_config")
    if waves_per_simd is not None or refill_below is not None or sort_rays is not None:
        cfg.waves_per_simd = waves_per_simd or cfg.waves_per_simd
        cfg.refill_below = refill_below or cfg.refill_below
        if sort_rays is not None:
            cfg.sort_rays = 1 if sort_rays else 0
        _check(LIB.vx_query_set_config(ctx._h, C.byref(cfg)), "vx_query_set_config")
    return cfg


def query_tune(ctx: "Context", scene_cam: Camera, rays: DeviceArray, hits: DeviceArray, n: int) -> QueryConfig:
    cfg = QueryConfig()
    _check(LIB.vx_query_tune(ctx._h, C.byref(scene_cam), rays.ptr, n, hits.ptr, C.byref(cfg)), "vx_query_tune")
    return cfg


_LIVE_SERVERS: "weakref.WeakSet[Server]" = weakref.WeakSet()


@atexit.register
def _close_live_servers() -> None:
    # A server's device thread uses its context: stop every server before
    # interpreter teardown can destroy contexts in arbitrary order.
    for srv in list(_LIVE_SERVERS):
        try:
            srv.close()
        except Exception:
            pass


class Server:
    """The serving loop (vx_server_*): continuous batching of view-stream frames
    and ray batches over one context, with admission and an edit stream."""

    def __init__(self, ctx: "Context", world: "HostWorld", memory_budget: int, max_views_per_launch: int = 64,
                 edit_threads: int = 8, max_queued_frames: int = 0, latency_target_ms: float = 0.0,
                 device_build: bool = False):
        cfg = ServerConfig()
        cfg.flags = SERVER_DEVICE_BUILD if device_build else 0
        cfg.memory_budget = memory_budget
        cfg.max_views_per_launch = max_views_per_launch
        cfg.edit_threads = edit_threads
        cfg.max_queued_frames = max_queued_frames
        cfg.latency_target_ms = latency_target_ms
        cfg.world = world.params
        for i in range(3):
            cfg.grid_origin[i] = world.grid_origin[i]
            cfg.grid_dims[i] = world.grid_dims[i]
        view = HostWorldView()
        _check(LIB.vx_host_world_get(world._h, C.byref(view)), "vx_host_world_get")
        h = C.c_void_p()
        _check(LIB.vx_server_create(ctx._h, C.byref(view), C.byref(cfg), C.byref(h)), "vx_server_create")
        self._h = h
        _LIVE_SERVERS.add(self)
        self.ctx, self.world = ctx, world
        self._keep = {}  # tag -> host buffers the server writes into

    def submit_view(self, stream_id: int, camera: Camera, width: int, height: int, tag: int,
                    out: np.ndarray | None = None) -> bool:
        """Queues one frame of a view stream; False when admission rejects it."""
        r = Request()
        r.kind, r.stream_id, r.width, r.height, r.camera, r.tag = REQ_VIEW, stream_id, width, height, camera, tag
        if out is not None:
            assert out.shape == (height, width, 4) and out.dtype == np.float32 and out.flags.c_contiguous
            r.out_rgba = out.ctypes.data
        return self._submit(r, tag, out)

    def submit_rays(self, camera: Camera, rays: np.ndarray, tag: int, out: np.ndarray | None = None) -> bool:
        rays = np.ascontiguousarray(rays, np.float32)
        r = Request()
        r.kind, r.camera, r.rays, r.n_rays, r.tag = REQ_RAYS, camera, rays.ctypes.data, len(rays), tag
        if out is not None:
            assert len(out) == len(rays) and out.dtype == HIT_DTYPE and out.flags.c_contiguous
            r.out_hits = out.ctypes.data
        return self._submit(r, tag, out)

    def _submit(self, r, tag, out) -> bool:
        rc = LIB.vx_server_submit(self._h, C.byref(r))
        if rc == VX_E_BUSY:
            return False
        _check(rc, "vx_server_submit")
        if out is not None:
            self._keep[tag] = out
        return True

    def poll(self) -> Result | None:
        res = Result()
        rc = LIB.vx_server_poll(self._h, C.byref(res))
        if rc == VX_E_AGAIN:
            return None
        _check(rc, "vx_server_poll")
        self._keep.pop(res.tag, None)
        return res

    def edit(self, xyz: np.ndarray, mats: np.ndarray) -> None:
        xyz = np.ascontiguousarray(xyz, np.int32).reshape(-1, 3)
        mats = np.ascontiguousarray(mats, np.uint8).reshape(-1)
        _check(LIB.vx_server_edit(self._h, xyz.ctypes.data, mats.ctypes.data, len(mats)), "vx_server_edit")

    def wait_edits(self) -> None:
        _check(LIB.vx_server_wait_edits(self._h), "vx_server_wait_edits")

    def drain(self) -> None:
        _check(LIB.vx_server_drain(self._h), "vx_server_drain")

    def close_stream(self, stream_id: int) -> None:
        _check(LIB.vx_server_close_stream(self._h, stream_id), "vx_server_close_stream")

    def stats(self) -> dict:
        st = ServerStats()
        _check(LIB.vx_server_get_stats(self._h, C.byref(st)), "vx_server_get_stats")
        return st.as_dict()

    def close(self) -> None:
        if self._h:
            LIB.vx_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Views of one context rendered together by fused launches (vx_batch_*)."""

    def __init__(self, ctx: Context, views):
        self.ctx = ctx
        self.views = list(views)  # keeps the views alive while the batch refers to them
        arr = (C.c_void_p * len(self.views))(*[v._h.value for v in self.views])
        h = C.c_void_p()
        _check(LIB.vx_batch_create(ctx._h, arr, len(self.views), C.byref(h)), "vx_batch_create")
        self._h = h

    def render(self, mask: int = PASS_FRAME, stream: int | None = None) -> None:
        _check(LIB.vx_batch_render(self._h, mask, stream), "vx_batch_render")

    def timing(self) -> float:
        ms = C.c_float()
        _check(LIB.vx_batch_timing(self._h, C.byref(ms)), "vx_batch_timing")
        return float(ms.value)

    def config(self) -> BatchConfig:
        cfg = BatchConfig()
        _check(LIB.vx_batch_get_config(self._h, C.byref(cfg)), "vx_batch_get_config")
        return cfg

    def set_config(self, waves_per_simd: int = 6, grab: int = 1, use_graph: bool = True, tile_shape: int = 0) -> None:
        cfg = BatchConfig(waves_per_simd, grab, 1 if use_graph else 0, tile_shape)
        _check(LIB.vx_batch_set_config(self._h, C.byref(cfg)), "vx_batch_set_config")

    def tune(self, mask: int = PASS_FRAME, iters: int = 2) -> BatchConfig:
        cfg = BatchConfig()
        _check(LIB.vx_batch_tune(self._h, mask, iters, C.byref(cfg)), "vx_batch_tune")
        return cfg

    def set_rays(self, scene_cam: Camera | None, rays: "DeviceArray | None", hits: "DeviceArray | None",
                 n: int) -> None:
        """vx_batch_set_rays: later renders with the primary pass also trace
        these n device-resident rays in the same launch (n = 0 detaches).
        The arrays are kept referenced until they are replaced."""
        self._rays = (rays, hits) if n else None
        _check(LIB.vx_batch_set_rays(self._h, C.byref(scene_cam) if scene_cam is not None else None,
                                     rays.ptr if n else None, n, hits.ptr if n else None), "vx_batch_set_rays")

    def close(self) -> None:
        if self._h:
            LIB.vx_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def algorithmic_bytes(counts: dict) -> int:
    """Bytes of the reference's loads/stores from a counter dict (struct sizes)."""
    return int(sum(counts[n] * b for n, b in zip(COUNT_NAMES, COUNT_BYTES)))
