"""Seeded workloads: worlds, camera poses and flights of the benchmark configs.

  * ``c0``: plain primary pass, 640x360, one camera at a fixed seeded pose
    (FOV 70 deg, pitch -15 deg) over an 8x4x8-chunk world (seed 42).
  * ``c1``: full pipeline over a 16x4x16-chunk world with a 5-level clipmap
    (256^2 texels per level, cell 1, 2, 4, 8, 16 m) and a 120-frame flight:
    a Catmull-Rom spline through 8 seeded waypoints traversed at 10 m/s,
    60 frames per second (1/6 m per frame).
All randomness comes from numpy's PCG64 seeded generators.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import (Camera, HostWorld, Pose, WorldParams, camera_build, clipmap_build, ground_height, make_pose,
               world_params)

VOXEL_M = 0.1
CHUNK_M = 6.4


@dataclass
class Scene:
    name: str
    params: WorldParams
    grid_origin: tuple
    grid_dims: tuple
    present: tuple
    render: tuple
    clip_levels: int = 5
    poses: list = field(default_factory=list)  # list of (Pose, time_s)
    stream_poses: list = field(default_factory=list)  # per client stream (serving), lists of (Pose, time_s)


def surface_m(params: WorldParams, x_m: float, z_m: float) -> float:
    return ground_height(params, int(math.floor(x_m / VOXEL_M)), int(math.floor(z_m / VOXEL_M))) * VOXEL_M


def seeded_pose(params: WorldParams, dims, seed: int, pitch_deg: float = -15.0, fov_deg: float = 70.0) -> Pose:
    rng = np.random.default_rng(seed)
    ext_x, ext_z = dims[0] * CHUNK_M, dims[2] * CHUNK_M
    x = float(rng.uniform(0.2 * ext_x, 0.8 * ext_x))
    z = float(rng.uniform(0.2 * ext_z, 0.8 * ext_z))
    y = surface_m(params, x, z) + float(rng.uniform(2.0, 6.0))
    yaw = float(rng.uniform(0.0, 2.0 * math.pi))
    return make_pose((x, y, z), yaw, math.radians(pitch_deg), fov_deg)


def catmull_rom(p0, p1, p2, p3, t):
    t2, t3 = t * t, t * t * t
    return 0.5 * ((2 * p1) + (-p0 + p2) * t + (2 * p0 - 5 * p1 + 4 * p2 - p3) * t2 + (-p0 + 3 * p1 - 3 * p2 + p3) * t3)


def flight_poses(params: WorldParams, dims, seed: int, frames: int, speed_mps: float = 10.0, fps: float = 60.0,
                 pitch_deg: float = -15.0, fov_deg: float = 70.0):
    """Poses along a spline through 8 seeded waypoints, arc-length stepped."""
    rng = np.random.default_rng(seed)
    ext_x, ext_z = dims[0] * CHUNK_M, dims[2] * CHUNK_M
    wps = []
    for _ in range(8):
        x = float(rng.uniform(0.15 * ext_x, 0.85 * ext_x))
        z = float(rng.uniform(0.15 * ext_z, 0.85 * ext_z))
        wps.append(np.array([x, surface_m(params, x, z) + float(rng.uniform(3.0, 8.0)), z]))
    pts = [wps[0]] + wps + [wps[-1]]
    dense = []
    for i in range(1, len(pts) - 2):
        for s in np.linspace(0.0, 1.0, 200, endpoint=False):
            dense.append(catmull_rom(pts[i - 1], pts[i], pts[i + 1], pts[i + 2], s))
    dense.append(pts[-2])
    dense = np.array(dense)
    seg = np.linalg.norm(np.diff(dense, axis=0), axis=1)
    arc = np.concatenate([[0.0], np.cumsum(seg)])
    step = speed_mps / fps
    out = []
    for f in range(frames):
        s = min(f * step, arc[-1] - 1e-3)
        i = int(np.searchsorted(arc, s, side="right") - 1)
        i = min(max(i, 0), len(dense) - 2)
        u = (s - arc[i]) / max(seg[i], 1e-9)
        p = dense[i] * (1 - u) + dense[i + 1] * u
        d = dense[i + 1] - dense[i]
        yaw = math.atan2(d[0], d[2])
        p[1] = max(p[1], surface_m(params, p[0], p[2]) + 1.5)
        out.append((make_pose(tuple(p), yaw, math.radians(pitch_deg), fov_deg), f / fps))
    return out


def c0(seed: int = 42) -> Scene:
    p = world_params(seed)
    dims = (8, 4, 8)
    s = Scene("c0_primary_640x360_8x4x8", p, (0, 0, 0), dims, (640, 360), (640, 360))
    s.poses = [(seeded_pose(p, dims, seed), 0.0)]
    return s


def c1(seed: int = 42, frames: int = 120, resolution=(1920, 1080)) -> Scene:
    p = world_params(seed)
    dims = (16, 4, 16)
    s = Scene(f"c1_pipeline_{resolution[0]}x{resolution[1]}_16x4x16_flight", p, (0, 0, 0), dims, tuple(resolution),
              tuple(resolution))
    s.poses = flight_poses(p, dims, seed, frames)
    return s


def frame_camera(scene: Scene, world: HostWorld, frame: int, prev_view_proj=None, *, wind_strength: float = 2.0,
                 time_s: float | None = None, jitter: bool = False, render=None, present=None) -> Camera:
    """Camera of `frame` of the scene's flight, at the scene's resolution unless
    render (and present, default = render) are given."""
    pose, t = scene.poses[frame % len(scene.poses)]
    render = render or scene.render
    present = present or (render if render is not scene.render else scene.present)
    return camera_build(pose, render, present, frame, time_s=t if time_s is None else time_s,
                        wind_strength=wind_strength, fog_density=0.002, grid_origin=scene.grid_origin,
                        grid_dims=scene.grid_dims, chunk_count=world.chunk_count, prev_view_proj=prev_view_proj,
                        jitter=jitter)


def scene_clipmap(scene: Scene, frame: int = 0):
    pose, _ = scene.poses[frame % len(scene.poses)]
    return clipmap_build(scene.params, pose.pos[0], pose.pos[2], levels=scene.clip_levels)


def query_rays(rng, n: int, grid_origin, grid_dims, tmax_m: float = 51.2) -> np.ndarray:
    """Ray-query workload rays: origins uniform in the world AABB, directions
    uniform on the unit sphere (normalised Gaussian triplets, renormalised in
    fp32), max distance tmax_m; an (n, 8) float32 array laid out as vx_ray."""
    lo = np.array(grid_origin, np.float64) * CHUNK_M
    hi = lo + np.array(grid_dims, np.float64) * CHUNK_M
    o = rng.uniform(lo, hi, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = o
    rays[:, 3] = tmax_m
    rays[:, 4:7] = d
    rays[:, 4:7] /= np.linalg.norm(rays[:, 4:7], axis=1, keepdims=True)
    return rays


def c4(seed: int = 42, streams: int = 32, frames: int = 600) -> Scene:
    """Serving workload world: the seed-42 generator over 32x4x32 chunks with a
    5-level clipmap; `streams` client view streams, each on its own seeded
    camera flight (waypoints seeded by seed + stream index)."""
    p = world_params(seed)
    dims = (32, 4, 32)
    s = Scene("c4", p, (0, 0, 0), dims, (1920, 1080), (1920, 1080), 5)
    s.poses = flight_poses(p, dims, seed, frames)
    s.stream_poses = [flight_poses(p, dims, seed + 1 + i, frames) for i in range(streams)]
    return s


VIEW_SIZES = [(320, 180), (640, 360), (1280, 720), (1920, 1080), (3840, 2160)]


def batch_views(n_views: int, n_poses: int, seed: int = 7):
    """The multi-view batching workload's views: resolutions drawn uniformly
    from VIEW_SIZES and starting frames of the scene's flight, both seeded.
    Returns (resolutions, frames)."""
    rng = np.random.default_rng(seed)
    res = [VIEW_SIZES[i] for i in rng.integers(0, len(VIEW_SIZES), n_views)]
    frames = rng.integers(0, n_poses, n_views)
    return res, frames


def query_batches(n_batches: int, grid_origin, grid_dims, seed: int = 11, lo: float = 1e3, hi: float = 1e6):
    """The incoherent ray-query workload: n_batches batch sizes log-uniform in
    [lo, hi] and all their rays (query_rays) from one seeded generator.
    Returns (sizes, rays)."""
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(lo), np.log(hi), n_batches)).astype(np.int64)
    rays = query_rays(rng, int(sizes.sum()), grid_origin, grid_dims, tmax_m=51.2)
    return sizes, rays
