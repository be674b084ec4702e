"""Shared test setup.

Markers
  gpu  -- needs an MI355X; runs the HIP path through the C ABI / pybind11 modules
          and compares with the CPU oracle (oracle/).  Deselect with -m "not gpu".
"""

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (HIP path vs the CPU oracle)")
    config.addinivalue_line("markers", "slow: long-running")
    config.addinivalue_line("markers", "cpu_only: passes placeholder pointers; skipped where a "
                            "GPU is visible")


@pytest.fixture(scope="session")
def golden_fchl19():
    return np.load(os.path.join(ROOT, "tests", "golden", "fchl19_ref.npz"))


@pytest.fixture(scope="session")
def config0():
    """Config-0 molecules (1,000, seed 0), padded."""
    from kernelforge_amd import synthetic

    coords, zs, energies = synthetic.random_molecules(1000, 0)
    C, Q, N = synthetic.pad_batch(coords, zs)
    return coords, zs, energies, C, Q, N
