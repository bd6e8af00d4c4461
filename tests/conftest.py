from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: long-running statistical check")


def batch_mean_error(x, n_batches: int = 64) -> tuple[float, float]:
    """64 batch means, as the reference test suite does (critmc tests/conftest.py:6-17)."""
    x = np.asarray(x, dtype=np.float64)
    size = len(x) // n_batches
    b = x[len(x) - size * n_batches :].reshape(n_batches, size).mean(axis=1)
    return float(x.mean()), float(b.std(ddof=1) / np.sqrt(n_batches))


def assert_consistent(sample, exact: float, z_max: float = 4.5, n_batches: int = 64):
    mean, err = batch_mean_error(sample, n_batches)
    z = (mean - exact) / err if err > 0 else 0.0
    assert abs(z) < z_max, f"mean {mean:.6f} +- {err:.2e} vs exact {exact:.6f} (z = {z:.2f})"


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle

    oracle.build()
    return oracle
