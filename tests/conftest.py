"""Shared pytest configuration: the `gpu` marker and common fixtures."""

import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

GOLDEN = ROOT / "tests" / "golden" / "reference_vectors.npz"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: CPU test taking more than ~30 s")


@pytest.fixture(scope="session")
def golden():
    return np.load(GOLDEN)


@pytest.fixture(scope="session")
def small_monitor(golden):
    """The reference-generated 19x19x23 H2 development monitor grid (fixture coordinates)."""
    from oracle import grid

    coords = np.stack([golden["mon_x"], golden["mon_y"], golden["mon_z"]], -1)
    return grid.monitor_geometry_from_coordinates(coords, tuple(map(tuple, golden["mon_box"])),
                                                  tuple(golden["mon_patch_radii"]))
