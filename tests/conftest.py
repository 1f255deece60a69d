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


@functools.lru_cache(maxsize=4)
def host_world(name: str):
    import voxhip
    from voxhip import scenes

    sc = {"c0": scenes.c0, "c1": scenes.c1}[name]()
    return sc, voxhip.HostWorld(sc.params, sc.grid_origin, sc.grid_dims)
