"""Pins for the oracle's spin-weighted transforms (Spin.jl), on the CPU: Wigner-d closed forms,
endpoints and symmetries (test_spin.jl:1-47), mpmath values from Wigner's explicit sum at high
degree (tests/golden/wigner_d.json), the spin ring engine against single-value wigner_d, exact poles
and the reflection identity (test_spin.jl:49-137), spin 0 = the scalar harmonic, the spin Direct
leaf against explicit sums of sYlm, and the spin stretched leaf against spin Direct (test_spin.jl:
139-175), with the adjoint identity throughout."""
import json
import math
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def Nl(l):
    return math.sqrt((2 * l + 1) / (4 * math.pi))


def test_wigner_closed_forms(O):
    b = 0.37
    c, s = math.cos(b), math.sin(b)
    cases = [((1, 0, 0), c), ((1, 1, 0), -s / math.sqrt(2)), ((1, 0, 1), s / math.sqrt(2)),
             ((1, 1, 1), (1 + c) / 2), ((1, 1, -1), (1 - c) / 2), ((1, -1, 0), s / math.sqrt(2)),
             ((1, 0, -1), -s / math.sqrt(2)), ((2, 0, 0), (3 * c * c - 1) / 2),
             ((3, 0, 0), (5 * c ** 3 - 3 * c) / 2),
             # l = 2 entries (standard table): d^2_{22}, d^2_{21}, d^2_{20}, d^2_{1,-1}
             ((2, 2, 2), ((1 + c) / 2) ** 2), ((2, 2, 1), -(1 + c) * s / 2),
             ((2, 2, 0), math.sqrt(6) / 4 * s * s)]
    for (l, m, n), ref in cases:
        assert O.wigner_d(l, m, n, b) == pytest.approx(ref, abs=1e-15), (l, m, n)
    # d^2_{1,-1} = (1 - cos b)(2 cos b + 1) / 2
    assert O.wigner_d(2, 1, -1, b) == pytest.approx((1 - c) * (2 * c + 1) / 2, abs=1e-15)


def test_wigner_endpoints_and_symmetry(O):
    assert O.wigner_d(5, 2, 2, 0.0) == 1.0 and O.wigner_d(5, 2, 3, 0.0) == 0.0
    assert abs(O.wigner_d(4, 2, -2, math.pi)) == 1.0
    for l, m, n in ((7, 3, -2), (40, 11, 5), (200, 60, -7)):
        a, b = O.wigner_d(l, m, n, 0.9), O.wigner_d(l, n, m, 0.9)
        assert a == pytest.approx((-1) ** (m - n) * b, abs=1e-12)
        # d^l_{mn} = d^l_{-n,-m} and d^l_{mn}(pi - b) = (-1)^(l+m) d^l_{m,-n}(b)
        assert a == pytest.approx(O.wigner_d(l, -n, -m, 0.9), abs=1e-12)
        assert O.wigner_d(l, m, n, math.pi - 0.9) == pytest.approx((-1) ** (l + m) * O.wigner_d(l, m, -n, 0.9),
                                                                     abs=1e-12)


def test_wigner_vs_mpmath(O):
    with open(os.path.join(HERE, "golden", "wigner_d.json")) as fh:
        cases = json.load(fh)
    for c in cases:
        g = O.wigner_d(c["l"], c["m"], c["n"], c["beta"])
        r = float(c["d"])
        assert abs(g) <= 1 + 1e-10
        assert abs(g - r) <= 1e-11 * abs(r) + 1e-300, c  # reference bound 1e-10 (test_spin.jl:30)


@pytest.mark.parametrize("L", [128, 256])
def test_spin_engine_vs_wigner(O, L):
    th = np.linspace(0.01, 0.6, 120)
    eng = O.SpinRings(th, L, 2, 0.0)
    for m in (L // 4, L // 2, 3 * L // 4, L - 1, -L // 2):
        a = np.zeros((L + 1, 2 * L + 1), complex)
        a[O.spin_coeff_index(L, m, L)] = 1
        G = eng.rings_from_alm(a)[:, m + L, 0]
        ref = np.array([Nl(L) * O.wigner_d(L, m, -2, t) for t in th])
        peak = np.abs(ref).max()
        assert peak > 0
        assert np.abs(G.real - ref).max() / peak < 1e-12 and np.all(G.imag == 0)


def test_spin_engine_poles_exact_and_reflection(O):
    L, s = 6, 2
    th = np.array([0.0, 0.3, math.pi / 2, math.pi - 0.3, math.pi])
    eng = O.SpinRings(th, L, s, 0.0)
    for m in (-3, 0, 2, 5):
        a = np.zeros((L + 1, 2 * L + 1), complex)
        a[O.spin_coeff_index(5, m, L)] = 1
        G = eng.rings_from_alm(a)[:, m + L, 0]
        ref = np.array([Nl(5) * O.wigner_d(5, m, -s, t) for t in th])
        assert np.abs(G.real - ref).max() < 1e-13
    L, s = 8, 1
    th = np.linspace(0.2, 1.1, 9)
    eng = O.SpinRings(th, L, s, 0.0, reflect=True)
    for l, m in ((3, 2), (7, -5), (8, 0)):
        a = np.zeros((L + 1, 2 * L + 1), complex)
        a[O.spin_coeff_index(l, m, L)] = 1
        G = eng.rings_from_alm(a)[:, m + L, 0]
        ref = np.array([Nl(l) * O.wigner_d(l, m, -s, math.pi - t) for t in th])
        assert np.abs(G.real - ref).max() < 1e-12


def test_spin_engine_adjoint_is_transpose(O, rng):
    L, s = 40, -2
    th = np.sort(rng.uniform(0.05, 1.2, 33))
    eng = O.SpinRings(th, L, s, 1e-10)
    a = rng.standard_normal((L + 1, 2 * L + 1)) + 1j * rng.standard_normal((L + 1, 2 * L + 1))
    a[~O.spin_valid_mask(L, s)] = 0
    G = rng.standard_normal((len(th), 2 * L + 1)) + 1j * rng.standard_normal((len(th), 2 * L + 1))
    Ga = eng.rings_from_alm(a)[:, :, 0]
    aG = eng.alm_from_rings(G)
    assert abs(np.vdot(G, Ga) - np.vdot(aG, a)) < 1e-12 * abs(np.vdot(G, Ga))
    assert np.all(aG[~O.spin_valid_mask(L, s)] == 0)


def test_sylm_spin0_is_ylm(O):
    from scipy.special import sph_harm_y
    th, ph = 0.8, 2.1
    for l, m in ((0, 0), (3, 2), (12, -5)):
        assert O.sylm(0, l, m, th, ph) == pytest.approx(complex(sph_harm_y(l, m, th, ph)), abs=1e-13)


def test_spin_direct_vs_sylm(O, rng):
    # test_spin.jl:146-158
    L, s, n = 8, 1, 25
    th = 0.4 * rng.uniform(size=n) + 0.2
    ph = 2 * math.pi * rng.uniform(size=n)
    p = O.Plan(th, ph, L, method="direct", spin=s)
    a = np.zeros((L + 1, 2 * L + 1), complex)
    a[O.spin_coeff_index(4, 2, L)] = 1
    ref = np.array([O.sylm(s, 4, 2, t, q) for t, q in zip(th, ph)])
    assert np.abs(p.synth(a) - ref).max() < 1e-12
    a = rng.standard_normal((L + 1, 2 * L + 1)) + 1j * rng.standard_normal((L + 1, 2 * L + 1))
    f = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    lhs = np.vdot(f, p.synth(a))
    assert abs(lhs - np.vdot(p.adjoint(f), a)) < 1e-12 * abs(lhs)
    # a random field of explicit sums over every valid mode
    a[~O.spin_valid_mask(L, s)] = 0
    f2 = np.array([sum(a[l, m + L] * O.sylm(s, l, m, t, q) for l in range(L + 1) for m in range(-l, l + 1))
                   for t, q in zip(th, ph)])
    assert np.abs(p.synth(a) - f2).max() < 1e-12 * np.abs(f2).max()


@pytest.mark.parametrize("s", [1, 2, -1])
@pytest.mark.parametrize("where", ["north", "south", "interior"])
def test_spin_stretch_vs_direct(O, rng, s, where):
    # test_spin.jl:160-175 (north cap), and the same in a southern cap (reflected plan) and an
    # interior cap (no DFS doubling)
    L, tol, n = 256, 1e-8, 120
    th = np.deg2rad(20) * np.sqrt(rng.uniform(size=n))
    if where == "south":
        th = math.pi - th
    elif where == "interior":
        th = 1.0 + 0.3 * rng.uniform(size=n)
    ph = 2 * math.pi * rng.uniform(size=n)
    pd = O.Plan(th, ph, L, method="direct", spin=s, nthreads=8)
    ps = O.Plan(th, ph, L, tol=tol, method="stretch", spin=s, nthreads=8)
    assert ps.info["kind_name"] == ("interior" if where == "interior" else "pole")
    assert ps.info["flip"] == (where == "south")
    a = rng.standard_normal((L + 1, 2 * L + 1)) + 1j * rng.standard_normal((L + 1, 2 * L + 1))
    fd, fs = pd.synth(a), ps.synth(a)
    assert np.abs(fs - fd).max() / np.abs(fd).max() < 10 * tol
    f = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    lhs = np.vdot(f, fs)
    assert abs(lhs - np.vdot(ps.adjoint(f), a)) < 1e-12 * abs(lhs)
    m = O.spin_valid_mask(L, s)
    ad, as_ = pd.adjoint(f), ps.adjoint(f)
    assert np.abs(as_[m] - ad[m]).max() / np.abs(ad[m]).max() < 10 * tol
