"""Shared pytest setup: import paths, the `gpu` marker, cached worlds."""
from __future__ import annotations

import functools
import sys
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parent.parent
for p in (REPO, REPO / "voxel-engine-v0_amd"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def vx():
    import voxhip

    return voxhip


@pytest.fixture(scope="session")
def gpu(vx):
    """The HIP library on a real device; fails (does not skip) if the library is missing."""
    if vx.device_count() == 0:
        pytest.skip("no GPU in this environment")
    return vx


def place_lights(sc, hw, n: int = 40, seed: int = 3) -> None:
    """LIGHT voxels one voxel above the ground at seeded points 2..25 m in
    front of the scene's first camera pose (the generated terrain has none)."""
    import math

    import numpy as np
    from voxhip import ground_height
    from voxhip.scenes import VOXEL_M

    rng = np.random.default_rng(seed)
    pose = sc.poses[0][0]
    fx, fz = math.sin(pose.yaw), math.cos(pose.yaw)
    xyz = []
    for _ in range(n):
        d = float(rng.uniform(2.0, 25.0))
        s = float(rng.uniform(-0.6, 0.6)) * d
        x, z = pose.pos[0] + fx * d + fz * s, pose.pos[2] + fz * d - fx * s
        wx, wz = int(math.floor(x / VOXEL_M)), int(math.floor(z / VOXEL_M))
        xyz.append((wx, ground_height(sc.params, wx, wz) + 1, wz))
    hw.edit(np.array(xyz, np.int32), np.full(len(xyz), 6, np.uint8))  # 6 = LIGHT


@functools.lru_cache(maxsize=4)
def host_world(name: str):
    """Cached worlds: "c0", "c1", and "c0_lit" (c0 with LIGHT voxels in view)."""
    import voxhip
    from voxhip import scenes

    sc = {"c0": scenes.c0, "c1": scenes.c1, "c0_lit": scenes.c0}[name]()
    hw = voxhip.HostWorld(sc.params, sc.grid_origin, sc.grid_dims)
    if name == "c0_lit":
        place_lights(sc, hw)
    return sc, hw
