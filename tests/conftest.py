"""Shared test setup: import paths, the `gpu` marker, and fixtures.

Tests marked `gpu` run the HIP library on a device and compare with the CPU oracle; the rest
pin the oracle to golden vectors and mathematics, and check the C ABI without a device.
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("oracle", "partialfsht.jl_amd"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs libpfsht_hip kernels)")


@pytest.fixture(scope="session")
def O():
    import oracle
    oracle.build()
    return oracle


@pytest.fixture
def rng():
    return np.random.default_rng(20251203)


def cap_points(rng, deg, M):
    """Area-uniform points on a polar cap: cos(theta) ~ U[cos(theta*), 1], phi ~ U[0, 2pi)."""
    ts = np.deg2rad(deg)
    th = np.arccos(rng.uniform(np.cos(ts), 1.0, M))
    ph = rng.uniform(0.0, 2 * np.pi, M)
    return th, ph


def relmax(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.abs(a - b).max() / np.abs(b).max())
