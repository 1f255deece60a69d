"""Helpers that run the same frame through the HIP passes and the oracle and
compare them to the parity tolerances:
  * hit identity (kind, chunk slot, leaf node, voxel, material): bit-exact;
  * depth: |d_gpu - d_ref| <= 1e-5 * |d_ref|;
  * colour (primary colour and post-composite output): <= 1e-3 absolute.
"""
from __future__ import annotations

import numpy as np

DEPTH_REL_TOL = 1e-5
COLOR_ABS_TOL = 1e-3


def make_context(vx, hw, clip_levels=5, godray_macro_fix=False):
    ctx = vx.Context(node_capacity=max(len(hw.nodes), 1), chunk_capacity=hw.chunk_count, grid_capacity=len(hw.grid),
                     clip_levels=clip_levels, clip_res=256, godray_macro_fix=godray_macro_fix)
    ctx.upload_world(hw)
    return ctx


def hit_mismatches(h_gpu, h_ref) -> np.ndarray:
    bad = h_gpu["kind"] != h_ref["kind"]
    bad |= h_gpu["mat"] != h_ref["mat"]
    bad |= h_gpu["slot"] != h_ref["slot"]
    bad |= h_gpu["node"] != h_ref["node"]
    bad |= np.any(h_gpu["voxel"] != h_ref["voxel"], axis=-1)
    return bad


def depth_rel_err(d_gpu, d_ref) -> float:
    return float(np.max(np.abs(d_gpu - d_ref) / np.maximum(np.abs(d_ref), 1e-30)))


def summary(h_gpu, h_ref, d_gpu, d_ref, c_gpu, c_ref) -> dict:
    mism = hit_mismatches(h_gpu, h_ref)
    return {
        "rays": int(mism.size),
        "identity_mismatches": int(mism.sum()),
        "t_bit_exact": bool(np.array_equal(h_gpu["t"].view(np.uint32), h_ref["t"].view(np.uint32))),
        "depth_max_rel": depth_rel_err(d_gpu, d_ref),
        "color_max_abs": float(np.max(np.abs(c_gpu - c_ref))),
        "color_bit_exact": bool(np.array_equal(c_gpu.view(np.uint32), c_ref.view(np.uint32))),
    }
