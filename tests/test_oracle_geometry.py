"""Pins the oracle's geometry, window and sizing (after src/Windows.jl, src/Stretch.jl:52-81).

Restates test/test_geometry_windows.jl:55-154; es_cdf is checked against mpmath quadrature of its
defining integral (tests/golden/es_cdf.json).
"""
import json
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_derived_relations(O):
    for (thhi, L, eps) in ((0.35, 512, 1e-8), (0.6, 1024, 1e-10), (1.0, 2048, 1e-6)):
        g = O.stretch_geometry(0.0, thhi, L, eps)
        beta = O.es_beta(eps)
        assert np.isclose(g.beta, beta, rtol=1e-15)
        assert np.isclose(g.epsapo, np.sqrt(2 * beta / (L * g.h)), rtol=1e-15)
        assert g.K == max(1, int(np.ceil(L * g.h * (1 + g.epsapo) ** 2 / np.pi)))
        assert np.isclose(g.Theta1, np.pi / (1 + g.epsapo), rtol=1e-15)
        assert np.isclose(g.sigma, (np.pi - g.Theta1) / 2, rtol=1e-15)
        assert np.isclose(g.ThetaT, g.Theta1 + g.sigma, rtol=1e-15)


def test_window_contract(O):
    g = O.stretch_geometry(0.0, 0.35, 512, 1e-8)
    for T in np.linspace(0, g.Theta1, 50):
        assert O.stretch_window(g, T) == 1.0
    assert O.stretch_window(g, np.pi) == 0.0
    ws = [O.stretch_window(g, T) for T in np.linspace(g.Theta1, np.pi, 200)]
    assert np.all(np.diff(ws) <= 1e-12)


def test_es_cdf_goldens(O):
    for e in json.load(open(os.path.join(GOLD, "es_cdf.json"))):
        ref = float(e["value"])
        assert abs(O.es_cdf(e["beta"], e["u"]) - ref) <= 1e-14 * max(1.0, abs(ref))
    beta = O.es_beta(1e-8)
    assert O.es_cdf(beta, -np.pi / 2) == 0.0
    us = np.linspace(-np.pi / 2, np.pi / 2, 60)
    v = [O.es_cdf(beta, u) for u in us]
    assert v == sorted(v)


def test_coordinate_maps_invert(O):
    for (lo, hi, L) in ((0.0, 0.4, 512), (0.8, 1.2, 1024), (0.0, np.pi, 256)):
        g = O.stretch_geometry(lo, hi, L, 1e-8)
        for th in np.linspace(max(lo, 1e-6), hi, 20):
            T = O.theta_to_stretch(g, th)
            assert abs(O.stretch_to_theta(g, T) - th) <= 1e-10


def test_interior_never_pads_across_pole(O):
    for (lo, hi, L) in ((0.02, 0.35, 256), (0.01, 0.5, 512), (2.7, 3.13, 512)):
        g = O.stretch_geometry(lo, hi, L, 1e-8)
        if g.kind == 1:
            pad = g.h * g.epsapo
            assert g.thlo - pad >= -1e-12 and g.thhi + pad <= np.pi + 1e-12


def test_degenerate_geometry_refused(O):
    assert O.stretch_geometry(0.0, 0.05, 32, 1e-10) is None
    assert O.stretch_geometry(0.4, 0.45, 16, 1e-12) is None
    with pytest.raises(ValueError):
        O.stretch_geometry(0.5, 0.4, 64, 1e-8)


def test_full_sphere(O):
    g = O.stretch_geometry(0.0, np.pi, 256, 1e-8)
    assert g.kind == 2 and g.K == 256
    assert O.stretch_window(g, 1.0) == 1.0


def test_smooth_sizes(O):
    def smooth(n):
        for p in (2, 3, 5):
            while n % p == 0:
                n //= p
        return n == 1
    for n in (1, 2, 7, 100, 122, 227, 1000, 4097):
        s = O.next_smooth(n)
        assert s >= n and smooth(s)
        e = O.next_smooth_even(n)
        assert e >= n and e % 2 == 0 and smooth(e)
        # minimality: no smaller 5-smooth number >= n
        assert all(not smooth(k) for k in range(n, s))
    assert O.next_smooth_even(135) == 144
    assert O.next_smooth_even(225) == 240


def test_config_sizes(O, rng):
    # plan sizing (src/Stretch.jl:164-194) at the benchmark configurations
    want = {(256, 20, 1e-8): (80, 160, 360, 178, 1613800),
            (2048, 20, 1e-8): (360, 720, 1920, 921, 304773641),
            (1024, 20, 1e-8): (216, 432, 1080, 513, 50996937),
            (2048, 30, 1e-5): (480, 960, 2430, 1214, 504108806)}
    for (L, deg, tol), (nh, nth, nph, mmax, W) in want.items():
        ts = np.deg2rad(deg)
        th = np.concatenate([[ts, 1e-3], np.arccos(rng.uniform(np.cos(ts), 1, 100))])
        ph = rng.uniform(0, 2 * np.pi, len(th))
        p = O.Plan(th, ph, L, tol=tol, method="stretch")
        i = p.info
        assert (i["nth_half"], i["nth"], i["nph"], i["mmax"], i["ring_work"]) == (nh, nth, nph, mmax, W)
        assert i["kind_name"] == "pole" and i["nph"] % 2 == 0


def test_auto_method(O, rng):
    # test_cost.jl:46-55: a cap picks Stretch, a tiny point set picks Direct
    th = 0.3 * np.sqrt(rng.uniform(size=3000))
    ph = rng.uniform(0, 2 * np.pi, 3000)
    assert O.Plan(th, ph, 256, tol=1e-8).method == "stretch"
    assert O.Plan(th[:10], ph[:10], 256, tol=1e-8).method == "direct"
