"""Shared test set-up: the `gpu` marker, import paths, grids and inputs."""
import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def gpu_available() -> bool:
    # the kernel driver's device node; present on a GPU host only
    return os.path.exists("/dev/kfd")


def pytest_collection_modifyitems(config, items):
    if gpu_available():
        return
    skip = pytest.mark.skip(reason="no GPU in this environment")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)


@pytest.fixture(scope="session")
def hip_lib():
    """The HIP library; it must be built and loadable on a GPU run."""
    from ntustcfdlab_amd import hip
    return hip.load_library()


def stretched_grid(nx, ny, nz):
    """A non-uniform grid, so that every metric index is exercised."""
    from ntustcfdlab_amd.grid import Grid, tanh_nodes
    return Grid(nx, ny, nz, 1.0, 1.3, 0.9,
                X=tanh_nodes(nx, 1.0, 0.4, 1.5), Y=tanh_nodes(ny, 1.3, 0.7, 1.2),
                Z=tanh_nodes(nz, 0.9, 0.45, 1.8))


def rand_halo(grid, rng, scale=1.0):
    a = grid.halo()
    a[...] = rng.uniform(-scale, scale, a.shape)
    return a


def rand_interior(grid, rng, lo=0.0, hi=1.0):
    a = grid.interior()
    a[...] = rng.uniform(lo, hi, a.shape)
    return a
