"""Shared fixtures.  GPU tests carry ``@pytest.mark.gpu``.

``pytest -m "not gpu"`` runs the CPU suite (oracle vs golden vectors, host
logic, C-ABI symbol export); ``pytest -m gpu`` runs the parity tests on an
MI355X through the HIP library.
"""

from __future__ import annotations

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built HIP library")


def load_golden(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def config2_layer_tuples():
    """9-layer profile of the spectral-frame workload: winds from seed=2.

    Speeds U(5, 35) m/s then directions U(0, 2 pi) rad, in that order; the
    direction is returned in degrees (the Layer convention).
    """
    rng = np.random.default_rng(2)
    speeds = rng.uniform(5.0, 35.0, 9)
    dirs = rng.uniform(0.0, 2.0 * np.pi, 9)
    heights = [0.0, 500.0, 1000.0, 2000.0, 4000.0, 8000.0, 12000.0, 16000.0, 20000.0]
    fracs = [0.5, 0.1, 0.08, 0.07, 0.06, 0.06, 0.05, 0.05, 0.03]
    return [(h, f, float(v), float(np.degrees(d)), 25.0)
            for h, f, v, d in zip(heights, fracs, speeds, dirs)]


@pytest.fixture(scope="session")
def hip():
    """The loaded HIP library (GPU tests only); fails loudly if absent."""
    import torch

    if not torch.cuda.is_available():
        pytest.fail("GPU test collected but no ROCm device is visible")
    from pyturb_amd import _hip

    return _hip.lib()
