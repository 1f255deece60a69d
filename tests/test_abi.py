"""The C-ABI library: it loads, exports every function include/voxhip.h
declares, reports errors through return codes and vx_last_error, and its
ctypes structs match the header layouts.  No GPU compute is called here."""
from __future__ import annotations

import ctypes as C
import re

import numpy as np
import pytest

from conftest import REPO


def _declared():
    text = (REPO / "include" / "voxhip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    return sorted(set(re.findall(r"\b(vx_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_function(vx):
    names = _declared()
    assert len(names) >= 30
    for n in names:
        assert hasattr(vx.LIB, n), f"{n} declared in include/voxhip.h but not exported"
    assert sorted(vx.EXPORTED) == names


def test_version_and_device_count(vx):
    assert vx.LIB.vx_version().decode().startswith("libvoxhip")
    assert vx.LIB.vx_device_count() >= 0


def test_struct_layouts(vx):
    assert C.sizeof(vx.Camera) == 464
    assert C.sizeof(vx.Clipmap) == 528
    assert C.sizeof(vx.Caps) == 32
    assert C.sizeof(vx.Region) == 24
    # CameraGpu field offsets (src/render/gpu_types.rs:52-86), also asserted in the C++ sources
    assert vx.Camera.cam_pos.offset == 256 and vx.Camera.chunk_size.offset == 320
    assert vx.Camera.voxel_params.offset == 336 and vx.Camera.grid_origin_chunk.offset == 352
    assert vx.Camera.render_present_px.offset == 384 and vx.Camera.profile_flags.offset == 400
    assert vx.Clipmap.level.offset == 16 and vx.Clipmap.offset.offset == 272


def test_create_without_device_fails_loudly(vx):
    if vx.LIB.vx_device_count() > 0:
        pytest.skip("a GPU is present")
    with pytest.raises(vx.VoxhipError, match="no GPU"):
        vx.Context(1024, 4, 4)


def test_null_arguments_are_rejected(vx):
    E_INVALID = -1
    L = vx.LIB
    assert L.vx_create(None, None) == E_INVALID
    assert "null" in L.vx_last_error().decode()
    assert L.vx_upload_regions(None, 0, None, 0) == E_INVALID
    assert L.vx_write_grid(None, None, 0) == E_INVALID
    assert L.vx_write_clipmap(None, None, None, 0) == E_INVALID
    assert L.vx_view_create(None, 16, 16, 16, 16, None) == E_INVALID
    assert L.vx_write_camera(None, None) == E_INVALID
    assert L.vx_render(None, 0x1F, None) == E_INVALID
    assert L.vx_read_output(None, 0, None, 0) == E_INVALID
    assert L.vx_view_counters(None, None, 0) == E_INVALID
    assert L.vx_view_timings(None, None) == E_INVALID
    assert L.vx_sync(None) == E_INVALID
    assert L.vx_build_chunk(None, None, 0, 4, None, None, 0, None, None, None) == E_INVALID
    L.vx_destroy(None)
    L.vx_view_destroy(None)


def test_build_chunk_rejects_small_output_capacity(vx):
    """A chunk whose tree exceeds the caller's buffers fails with E_CAPACITY."""
    p = vx.world_params(42)
    mat, ground = vx.fill_chunk(p, 3, 1, 2)
    full = vx.build_chunk(mat, ground, 64, p.dirt_depth)
    with pytest.raises(vx.VoxhipError):
        vx.build_chunk(mat, ground, 64, p.dirt_depth, node_capacity=len(full.nodes) - 1)
    again = vx.build_chunk(mat, ground, 64, p.dirt_depth, node_capacity=len(full.nodes))
    assert np.array_equal(again.nodes.view(np.uint32), full.nodes.view(np.uint32))
