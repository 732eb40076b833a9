"""Bounding box and the analysis restatements, on the CPU.

The library's pfsht_bounding_box (host code, no device work) against the oracle's sort-based
restatement of src/Box.jl:43-153 and closed forms; the oracle's analysis restatements
(src/Analysis.jl) pinned by properties the reference's own tests assert (test_analysis.jl)."""
import ctypes as C

import numpy as np
import pytest

from conftest import cap_points


def _box(theta, phi):
    import pfsht
    return pfsht.bounding_box(theta, phi)


def test_box_closed_forms():
    # gaps 0.1, 2.8 and the wrap 2 pi - 2.9: the arc starts at 0.1 and spans 2.9
    b = _box([0.3, 0.4, 0.5], [0.1, 0.2, 3.0])
    assert b["phistart"] == 0.1 and abs(b["phiext"] - 2.9) < 1e-15
    assert abs(b["fsky"] - (np.cos(0.3) - np.cos(0.5)) * 2.9 / (4 * np.pi)) < 1e-16
    assert abs(b["dl"] - np.pi / max(0.2, 2.9 * np.sin(0.4))) < 1e-14
    # a point on the pole: every longitude, whatever phi holds
    b = _box([0.0, 0.2], [1.0, 1.1])
    assert b["touches_north"] and b["phistart"] == 0.0 and b["phiext"] == 2 * np.pi
    b = _box([np.pi, 2.0], [1.0, 1.1])
    assert b["touches_south"] and b["phiext"] == 2 * np.pi
    # an arc across phi = 0 (longitudes given as negative angles and beyond 2 pi)
    b = _box([0.5, 0.6, 0.7], [-0.2, 0.1, 2 * np.pi + 0.05])
    assert abs(b["phistart"] - (2 * np.pi - 0.2)) < 1e-15 and abs(b["phiext"] - 0.3) < 1e-15
    # one point, and repeated longitudes
    b = _box([0.4], [7.0])
    assert b["phiext"] == 0.0 and b["fsky"] == 0.0 and b["dl"] == np.inf
    b = _box([0.4, 0.5], [1.0, 1.0])
    assert b["phiext"] == 0.0


@pytest.mark.parametrize("seed", range(6))
def test_box_vs_oracle(O, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(2, 5000))
    kind = seed % 3
    if kind == 0:      # a longitude arc
        th = rng.uniform(0.2, 1.4, n)
        ph = rng.uniform(-1.0, 2.5, n)
    elif kind == 1:    # clustered longitudes with a few outliers
        th = rng.uniform(1.0, 2.0, n)
        ph = np.concatenate([rng.normal(3.0, 0.05, n - 3), [0.5, 5.9, 6.2]])
    else:              # a polar cap without a point on the pole
        th, ph = cap_points(rng, 20, n)
    b, o = _box(th, ph), O.bounding_box(th, ph)
    for k in ("thlo", "thhi", "phistart", "phiext", "touches_north", "touches_south"):
        assert b[k] == o[k], k
    assert abs(b["fsky"] - o["fsky"]) <= 1e-15 * max(1.0, abs(o["fsky"]))
    assert abs(b["dl"] - o["dl"]) <= 1e-13 * abs(o["dl"])


def test_box_argument_errors():
    import pfsht
    L = pfsht.lib()
    out = pfsht.BoxC()
    assert L.pfsht_bounding_box(None, None, 3, C.byref(out)) == 2
    x = np.zeros(1)
    assert L.pfsht_bounding_box(x.ctypes.data, x.ctypes.data, 0, C.byref(out)) == 2


def test_oracle_solve_converges_on_covered_sphere(O):
    # test_analysis.jl:152-168: near-global coverage, lmax 6 — CG fits to 1e-9 and recovers the truth
    rng = np.random.default_rng(56)
    n, L = 3000, 6
    th = np.arccos(2 * rng.uniform(size=n) - 1)
    ph = 2 * np.pi * rng.uniform(size=n)
    plan = O.Plan(th, ph, L, method="direct")
    at = O.rand_alm_fast(L, rng)
    f = plan.synth(at)
    x, info = O.solve(plan, f, mode="cg", rtol=1e-12, maxiter=500)
    assert np.linalg.norm(plan.synth(x) - f) / np.linalg.norm(f) < 1e-9
    m = O.valid_mask(L)
    assert np.abs(x[m] - at[m]).max() / np.abs(at[m]).max() < 1e-8
    assert info["iters"] >= 1 and info["relres"] <= 1e-12


def test_oracle_pseudo_spectrum_parseval(O, rng):
    # sum_l (2l + 1) C_l = |A' (w f)|^2: every mode is counted once
    th, ph = cap_points(rng, 25, 600)
    L = 12
    plan = O.Plan(th, ph, L, method="direct")
    f = rng.standard_normal(600) + 1j * rng.standard_normal(600)
    w = rng.uniform(0.5, 1.5, 600)
    Cl = O.pseudo_spectrum(plan, f, w)
    a = plan.adjoint(np.asfortranarray(f * w))
    assert abs(np.sum((2 * np.arange(L + 1) + 1) * Cl) - np.sum(np.abs(a) ** 2)) < 1e-12 * np.sum(np.abs(a) ** 2)
