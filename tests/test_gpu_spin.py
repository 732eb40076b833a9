"""Spin-weighted transforms on the device (src/Spin.jl) through the C ABI: the spin ring engine
against the oracle's and against single-value wigner_d (test_spin.jl:49-137), the spin Direct leaf
against explicit sYlm (test_spin.jl:146-158), the spin stretched leaf (north, reflected south and
interior caps; several transforms) against the oracle's spin plan and spin Direct
(test_spin.jl:160-175), the multi-box leaf with spin, and the adjoint identity throughout.
Coefficients use the dense (l, m + lmax) layout. Tolerances are written in each test."""
import math
import os

import numpy as np
import pytest

from conftest import relmax

pytestmark = pytest.mark.gpu
NT = min(16, os.cpu_count() or 1)


@pytest.fixture(scope="module")
def torch():
    import torch as T
    assert T.cuda.is_available(), "a GPU is required for the gpu-marked tests"
    return T


@pytest.fixture(scope="module")
def P():
    import pfsht
    pfsht.lib()
    return pfsht


def _rel_valid(O, a, b, L, s):
    m = O.spin_valid_mask(L, s)
    return float(np.abs(a[m] - b[m]).max() / np.abs(b[m]).max())


# B = 2 and B = 5 cover the batched DPP forward with a group of two and with groups of four and one
# (the second launch at transform offset 4)
@pytest.mark.parametrize("L,s,reflect,B", [(128, 2, False, 1), (512, -1, False, 1), (256, 3, True, 1),
                                           (200, 2, False, 3), (160, 1, False, 2), (200, -2, True, 5)])
def test_spin_engine_vs_oracle(torch, P, O, rng, L, s, reflect, B):
    th = np.sort(0.02 + 0.9 * rng.uniform(size=150))
    eng = P.Rings(th, L, 1e-8, ntrans=B, reflect=reflect, spin=s)
    ref = O.SpinRings(th, L, s, 1e-8, ntrans=B, reflect=reflect)
    assert eng.mmax == ref.mmax and eng.work == ref.work
    a = rng.standard_normal((L + 1, 2 * L + 1, B)) + 1j * rng.standard_normal((L + 1, 2 * L + 1, B))
    a[~O.spin_valid_mask(L, s)] = 0
    a = np.asfortranarray(a)
    G = np.zeros((len(th), 2 * L + 1, B), complex, order="F")
    eng.rings_from_alm(G, a)
    Gr = ref.rings_from_alm(a)
    assert relmax(G, Gr) < 1e-12  # per-lane exponents and FMA contraction: round-off
    Gi = np.asfortranarray(rng.standard_normal(G.shape) + 1j * rng.standard_normal(G.shape))
    ad = np.zeros_like(a)
    eng.alm_from_rings(ad, Gi)
    ar = ref.alm_from_rings(Gi).reshape(a.shape, order="F")
    assert _rel_valid(O, ad, ar, L, s) < 1e-12
    assert np.all(ad[~O.spin_valid_mask(L, s)] == 0)
    # the device adjoint is deterministic
    ad2 = np.zeros_like(a)
    eng.alm_from_rings(ad2, Gi)
    assert np.array_equal(ad, ad2)


def test_spin_engine_vs_wigner_and_poles(torch, P, O):
    th = np.linspace(0.01, 0.6, 200)
    L = 256
    eng = P.Rings(th, L, 0.0, spin=2)
    for m in (L // 4, L // 2, 3 * L // 4, L - 1):
        a = np.zeros((L + 1, 2 * L + 1), complex)
        a[P.spin_coeff_index(L, m, L)] = 1
        G = np.zeros((len(th), 2 * L + 1), complex, order="F")
        eng.rings_from_alm(G, np.asfortranarray(a))
        ref = np.array([math.sqrt((2 * L + 1) / (4 * math.pi)) * O.wigner_d(L, m, -2, t) for t in th])
        assert np.abs(G[:, m + L].real - ref).max() / np.abs(ref).max() < 1e-12
    L, s = 6, 2
    th = np.array([0.0, 0.3, math.pi / 2, math.pi - 0.3, math.pi])
    eng = P.Rings(th, L, 0.0, spin=s)
    for m in (-3, 0, 2, 5):
        a = np.zeros((L + 1, 2 * L + 1), complex)
        a[P.spin_coeff_index(5, m, L)] = 1
        G = np.zeros((len(th), 2 * L + 1), complex, order="F")
        eng.rings_from_alm(G, np.asfortranarray(a))
        ref = np.array([math.sqrt(11 / (4 * math.pi)) * O.wigner_d(5, m, -s, t) for t in th])
        assert np.abs(G[:, m + L].real - ref).max() < 1e-13


def test_spin_direct_vs_sylm(torch, P, O, rng):
    L, s, n = 8, 1, 25
    th = 0.4 * rng.uniform(size=n) + 0.2
    ph = 2 * math.pi * rng.uniform(size=n)
    p = P.Plan(th, ph, lmax=L, method="direct", spin=s)
    assert p.info["spin"] == s and p.algorithm == "direct"
    a = np.zeros((L + 1, 2 * L + 1), complex, order="F")
    a[P.spin_coeff_index(4, 2, L)] = 1
    ref = np.array([O.sylm(s, 4, 2, t, q) for t, q in zip(th, ph)])
    assert np.abs((p @ a) - ref).max() < 1e-12
    a = np.asfortranarray(rng.standard_normal((L + 1, 2 * L + 1)) + 1j * rng.standard_normal((L + 1, 2 * L + 1)))
    f = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    lhs = np.vdot(f, p @ a)
    assert abs(lhs - np.vdot(p.H @ f, a)) < 1e-12 * abs(lhs)
    po = O.Plan(th, ph, L, method="direct", spin=s)
    assert relmax(p @ a, po.synth(a)) < 1e-12
    assert _rel_valid(O, p.H @ f, po.adjoint(f), L, s) < 1e-12


@pytest.mark.parametrize("s", [1, 2, -1])
@pytest.mark.parametrize("where", ["north", "south", "interior"])
def test_spin_stretch_vs_oracle_and_direct(torch, P, O, rng, s, where):
    L, tol, n = 256, 1e-8, 3000
    th = np.deg2rad(20) * np.sqrt(rng.uniform(size=n))
    if where == "south":
        th = math.pi - th
    elif where == "interior":
        th = 1.0 + 0.3 * rng.uniform(size=n)
    ph = 2 * math.pi * rng.uniform(size=n)
    pg = P.Plan(th, ph, lmax=L, tol=tol, method="stretch", spin=s)
    po = O.Plan(th, ph, L, tol=tol, method="stretch", spin=s, nthreads=NT)
    for k in ("nth", "nph", "mmax", "ring_work", "flip"):
        assert pg.info[k] == po.info[k], k
    a = np.asfortranarray(rng.standard_normal((L + 1, 2 * L + 1)) + 1j * rng.standard_normal((L + 1, 2 * L + 1)))
    fg = pg @ a
    assert relmax(fg, po.synth(a)) < 1e-11
    sub = rng.choice(n, 200, replace=False)
    pd = O.Plan(th[sub], ph[sub], L, method="direct", spin=s, nthreads=NT)
    assert relmax(fg[sub], pd.synth(a)) < 10 * tol  # test_spin.jl:172
    f = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    ag = pg.H @ f
    assert _rel_valid(O, ag, po.adjoint(f), L, s) < 1e-11
    assert np.all(ag[~O.spin_valid_mask(L, s)] == 0)
    lhs = np.vdot(f, fg)
    assert abs(lhs - np.vdot(ag, a)) / abs(lhs) < 1e-12  # test_spin.jl:173
    assert np.array_equal(pg.H @ f, ag)


def test_spin_ntrans_and_multibox(torch, P, O, rng):
    L, tol, B, s = 192, 1e-8, 3, 2
    n1 = 1500
    th = np.concatenate([0.3 * np.sqrt(rng.uniform(size=n1)), math.pi - 0.3 * np.sqrt(rng.uniform(size=n1))])
    ph = 2 * math.pi * rng.uniform(size=2 * n1)
    a = np.asfortranarray(rng.standard_normal((L + 1, 2 * L + 1, B)) + 1j * rng.standard_normal((L + 1, 2 * L + 1, B)))
    f = np.asfortranarray(rng.standard_normal((2 * n1, B)) + 1j * rng.standard_normal((2 * n1, B)))
    pg = P.Plan(th, ph, lmax=L, tol=tol, ntrans=B, method="multibox", spin=s)
    po = O.MultiBox(th, ph, L, tol=tol, ntrans=B, nthreads=NT, spin=s)
    assert pg.info["nboxes"] == len(po.groups) == 2
    assert relmax(pg @ a, po.synth(a)) < 1e-11
    ag = pg.H @ f
    assert _rel_valid(O, ag, po.adjoint(f), L, s) < 1e-11
    lhs = np.vdot(f, pg @ a)
    assert abs(lhs - np.vdot(ag, a)) / abs(lhs) < 1e-12
    # each column equals the single-transform plan
    p1 = P.Plan(th[:n1], ph[:n1], lmax=L, tol=tol, method="stretch", spin=s)
    pB = P.Plan(th[:n1], ph[:n1], lmax=L, tol=tol, ntrans=B, method="stretch", spin=s)
    fB = pB @ a
    aB = pB.H @ np.asfortranarray(f[:n1])
    for b in range(B):
        assert relmax(fB[:, b], p1 @ np.asfortranarray(a[:, :, b])) < 1e-13
        assert _rel_valid(O, aB[:, :, b], p1.H @ np.ascontiguousarray(f[:n1, b]), L, s) < 1e-13


def test_spin_plan_errors_and_analysis(torch, P, O, rng):
    th = np.deg2rad(20) * np.sqrt(rng.uniform(size=2000))
    ph = 2 * math.pi * rng.uniform(size=2000)
    with pytest.raises(P.ArgumentError):
        P.Plan(th, ph, lmax=4, tol=1e-8, spin=5)
    L = 128
    p = P.Plan(th, ph, lmax=L, tol=1e-8, method="stretch", spin=2)
    a = np.asfortranarray(rng.standard_normal((L + 1, 2 * L + 1)) + 1j * rng.standard_normal((L + 1, 2 * L + 1)))
    w = rng.uniform(0.5, 1.5, 2000)
    assert np.array_equal(p.apply_resolution(a, w), p.H @ np.asfortranarray((p @ a) * w))
    with pytest.raises(P.ArgumentError):
        p.pseudo_spectrum(p @ a, w)
    with pytest.raises(P.ArgumentError):
        p.solve(p @ a, mode="wiener", prior=np.ones(L + 1))
    x, info = p.solve(p @ a, mode="cg", rtol=1e-10, maxiter=5)
    assert info["iters"] == 5
