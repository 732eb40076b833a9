"""Pins the oracle's associated-Legendre engine (oracle/pfsht_oracle.c, after src/Legendre.jl).

Restates test/test_legendre.jl and test/test_backends.jl:202-238 of the reference. The checks
that pin values come from outside the recurrence: mpmath's hypergeometric Ferrers functions
(tests/golden/alf.json), scipy's spherical harmonics, and closed forms.
"""
import json
import os

import numpy as np
import pytest
from scipy.special import sph_harm_y

from conftest import relmax

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def ylm(O, l, m, th, ph):
    return O.alf(l, m, th) * np.exp(1j * m * ph)


def test_alf_vs_mpmath_goldens(O):
    # test_legendre.jl:2-27 (BigFloat at the turning point and fixed tuples), bound 1e-11
    worst = 0.0
    for e in json.load(open(os.path.join(GOLD, "alf.json"))):
        ref = float(e["value"])
        got = O.alf(e["l"], e["m"], e["theta"])
        if abs(ref) < 2.2250738585072014e-308:
            continue
        assert got != 0.0
        worst = max(worst, abs(got - ref) / abs(ref))
    assert worst < 1e-11, worst


def test_closed_form_ylm(O):
    # test_legendre.jl:29-36
    th, ph = 0.7, 1.1
    i4p = 1 / np.sqrt(4 * np.pi)
    assert np.isclose(ylm(O, 0, 0, 0.3, 1.2), i4p, rtol=1e-14)
    assert np.isclose(ylm(O, 1, 0, th, ph), np.sqrt(3 / (4 * np.pi)) * np.cos(th), rtol=1e-14)
    assert np.isclose(ylm(O, 1, 1, th, ph), -np.sqrt(3 / (8 * np.pi)) * np.sin(th) * np.exp(1j * ph), rtol=1e-14)
    assert np.isclose(ylm(O, 1, -1, th, ph), np.sqrt(3 / (8 * np.pi)) * np.sin(th) * np.exp(-1j * ph), rtol=1e-14)
    assert np.isclose(ylm(O, 2, 2, th, ph), np.sqrt(15 / (32 * np.pi)) * np.sin(th) ** 2 * np.exp(2j * ph),
                      rtol=1e-14)
    for (l, m) in ((3, 2), (7, 5), (40, 13)):  # Pbar_l^-m = (-1)^m Pbar_l^m
        assert np.isclose(O.alf(l, -m, th), (-1) ** m * O.alf(l, m, th), rtol=1e-14)


def test_alf_vs_scipy_sph_harm(O, rng):
    # scipy's Y_l^m (Condon-Shortley, orthonormal) is an independent evaluation
    for _ in range(200):
        l = int(rng.integers(0, 60))
        m = int(rng.integers(-l, l + 1))
        th = float(rng.uniform(0.01, np.pi - 0.01))
        ph = float(rng.uniform(0, 2 * np.pi))
        ref = sph_harm_y(l, m, th, ph)
        got = ylm(O, l, m, th, ph)
        assert abs(got - ref) <= 1e-12 * max(1.0, abs(ref)), (l, m, th)


def test_polar_truncation_drops_only_negligible(O):
    # test_legendre.jl:52-67
    for (L, th, eps) in ((512, 0.4, 1e-8), (1024, 0.6, 1e-8), (2048, 0.35, 1e-6)):
        mm = O.polar_mmax(th, L, eps)
        if mm >= L:
            continue
        peak = max(abs(O.alf(L, m, th)) for m in range(mm + 1))
        dropped = max(abs(O.alf(L, m, th)) for m in range(mm + 1, min(L, mm + 40) + 1))
        assert dropped <= eps * peak
    ms = [O.polar_mmax(t, 512, 1e-8) for t in np.linspace(1e-4, np.pi / 2, 40)]
    assert ms == sorted(ms)


def test_ring_engine_transpose(O, rng):
    # test_legendre.jl:69-85: <R a, g> = <a, R^T g> to round-off
    for L in (16, 64):
        th = np.sort(0.05 + 0.9 * rng.uniform(size=37))
        eng = O.Rings(th, L, 1e-12)
        a = rng.standard_normal((L + 1, 2 * L + 1)) + 1j * rng.standard_normal((L + 1, 2 * L + 1))
        G = eng.rings_from_alm(a)
        Gr = rng.standard_normal(G.shape) + 1j * rng.standard_normal(G.shape)
        ar = eng.alm_from_rings(Gr)
        lhs = np.vdot(G, Gr)
        rhs = np.vdot(a.reshape(ar.shape, order="F"), ar)
        assert abs(lhs - rhs) <= 1e-11 * abs(lhs)


def test_ring_engine_matches_per_point(O):
    # test_legendre.jl:87-98
    L = 48
    th = np.sort([0.02, 0.11, 0.37, 0.8, 1.4, 2.7])
    eng = O.Rings(th, L, 1e-13)
    a = np.zeros((L + 1, 2 * L + 1), complex, order="F")
    r, c = O.sph_mode(30, 7)
    a[r, c] = 1.0
    G = eng.rings_from_alm(a)
    for j, t in enumerate(th):
        assert abs(G[j, 7 + L, 0] - O.alf(30, 7, t)) <= 1e-13


def test_ring_engine_matches_scipy_sum(O, rng):
    # G[j, m] = sum_l a_lm Pbar_l^m(theta_j): an independent evaluation through scipy
    L = 40
    th = np.sort(rng.uniform(0.05, 3.0, 9))
    eng = O.Rings(th, L, 1e-14)
    a = O.rand_alm_fast(L, rng)
    G = eng.rings_from_alm(a)[:, :, 0]
    for m in (-17, -3, 0, 5, 21):
        ref = np.zeros(len(th), complex)
        for l in range(abs(m), L + 1):
            r, c = O.sph_mode(l, m)
            ref += a[r, c] * np.real(sph_harm_y(l, m, th, 0.0))
        assert relmax(G[:, m + L], ref) < 1e-12


def test_reflection_identity(O, rng):
    # reflect=True realises Y_lm(pi - theta) = (-1)^(l+m) Y_lm(theta) with one sign on cos theta
    # (src/Legendre.jl:215-228): the reflected engine on theta must equal the plain engine on pi - theta.
    L = 60
    th = np.sort(rng.uniform(0.1, 0.9, 11))
    a = O.rand_alm_fast(L, rng)
    Gr = O.Rings(th, L, 1e-12, reflect=True).rings_from_alm(a)
    Gp = O.Rings(np.sort(np.pi - th), L, 1e-12).rings_from_alm(a)
    assert relmax(Gr, Gp[::-1]) < 1e-12


def test_nested_ranges_match_scan(O):
    # test_legendre.jl:109-128
    for (lo, hi, L) in ((1e-3, 0.4, 256), (1.2, 2.0, 256), (0.5, 2.6, 512)):
        th = np.linspace(lo, hi, 77)
        mlim = np.array([O.polar_mmax(t, L, 1e-8) for t in th], np.int32)
        jlo, jhi = O.nested_ranges(mlim, L)
        for m in range(L + 1):
            keep = np.nonzero(mlim >= m)[0]
            if len(keep) == 0:
                assert jhi[m] < jlo[m]
            else:
                assert (jlo[m], jhi[m]) == (keep[0], keep[-1])
                assert list(keep) == list(range(keep[0], keep[-1] + 1))


def test_counted_work(O):
    # test_legendre.jl:100-107, test_cost.jl:16-26
    L = 512
    th = np.linspace(1e-3, 0.5, 120)
    mlim = [O.polar_mmax(t, L, 1e-8) for t in th]
    w_opt = O.retained_triples(mlim, L)
    assert 0 < w_opt < len(th) * sum(L - m + 1 for m in range(L + 1))
    for (L, hi) in ((128, 0.3), (512, 0.5), (512, 1.4)):
        th = np.linspace(1e-3, hi, 90)
        mlim = [O.polar_mmax(t, L, 1e-8) for t in th]
        eng = O.Rings(th, L, 1e-8)
        assert O.retained_triples(mlim, L) == eng.work
        assert max(mlim) == eng.mmax


def test_threaded_equals_serial_bitwise(O, rng):
    # test_backends.jl:202-238: interleaved orders per thread, each order summed identically
    for L in (64, 257):
        th = np.sort(0.01 + 0.5 * rng.uniform(size=53))
        a = O.rand_alm_fast(L, rng)
        eng = O.Rings(th, L, 1e-8)
        Gs = eng.rings_from_alm(a, nthreads=1)
        Gt = eng.rings_from_alm(a, nthreads=4)
        assert np.array_equal(Gs, Gt)
        assert np.array_equal(eng.alm_from_rings(Gs, 1), eng.alm_from_rings(Gs, 4))


def test_unsorted_rings_refused(O):
    with pytest.raises(ValueError):
        O.Rings(np.array([0.3, 0.1]), 8, 1e-8)
