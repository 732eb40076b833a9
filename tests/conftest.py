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


def centered_cap_points(rng, deg, M, theta_c, phi_c=0.0):
    """Area-uniform points on the cap of radius deg degrees around (theta_c, phi_c): a polar cap
    rotated about the y axis by theta_c, then about z by phi_c."""
    th, ph = cap_points(rng, deg, M)
    x, y, z = np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)
    c, s = np.cos(theta_c), np.sin(theta_c)
    x, z = c * x + s * z, -s * x + c * z
    return np.arccos(np.clip(z, -1.0, 1.0)), np.mod(np.arctan2(y, x) + phi_c, 2 * np.pi)


def relmax(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.abs(a - b).max() / np.abs(b).max())
