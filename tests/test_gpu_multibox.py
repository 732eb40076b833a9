"""Multi-box leaf on the device (src/MultiBox.jl) through the C ABI: a stretched plan per box of
the partition, synthesis as the disjoint union of the boxes' fields, adjoint as the sum of the
boxes' coefficients in box order. Against the oracle's multi-box restatement on the same inputs,
the FP64 Direct oracle (test_transforms.jl:210-227: 10 tol) and the adjoint identity (1e-12);
AUTO selection against the oracle's select_method (test_cost.jl:66-73). Tolerances are written in
each test."""
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


def _caps(rng, n1, r):
    th = np.concatenate([r * np.sqrt(rng.uniform(size=n1)), np.pi - r * np.sqrt(rng.uniform(size=n1))])
    return th, 2 * np.pi * rng.uniform(size=2 * n1)


def _rel_modes(O, a, b, L):
    m = O.valid_mask(L)
    return float(np.abs(a[m] - b[m]).max() / np.abs(b[m]).max())


def test_two_caps_vs_direct(torch, P, O):
    # test_transforms.jl:210-227 (n1 = 80 per cap, radius 0.25, lmax 256, tol 1e-8)
    rng = np.random.default_rng(37)
    L, tol = 256, 1e-8
    th, ph = _caps(rng, 80, 0.25)
    pg = P.Plan(th, ph, lmax=L, tol=tol, method="multibox")
    assert pg.algorithm == "multibox" and pg.info["nboxes"] == 2
    pd = O.Plan(th, ph, L, method="direct", nthreads=NT)
    a = O.rand_alm_fast(L, rng)
    fg = pg @ a
    assert relmax(fg, pd.synth(a)) < 10 * tol
    f = rng.standard_normal(len(th)) + 1j * rng.standard_normal(len(th))
    ag = pg.H @ f
    lhs = np.vdot(f, fg)
    assert abs(lhs - np.vdot(ag, a)) / abs(lhs) < 1e-12


@pytest.mark.parametrize("B", [1, 2])
def test_caps_vs_oracle_multibox(torch, P, O, B):
    rng = np.random.default_rng(38)
    L, tol = 256, 1e-8
    th, ph = _caps(rng, 3000, 0.3)
    pg = P.Plan(th, ph, lmax=L, tol=tol, ntrans=B, method="multibox")
    po = O.MultiBox(th, ph, L, tol=tol, ntrans=B, nthreads=NT)
    assert pg.info["nboxes"] == len(po.groups) == 2
    a = O.rand_alm_fast(L, rng) if B == 1 else np.stack([O.rand_alm_fast(L, rng) for _ in range(B)], axis=-1)
    a = np.asfortranarray(a)
    fg = pg @ a
    fo = po.synth(a)
    assert relmax(fg, fo) < 1e-11  # same boxes, same stretched plans per box: far inside tol
    sub = rng.choice(len(th), 300, replace=False)
    pd = O.Plan(th[sub], ph[sub], L, ntrans=B, method="direct", nthreads=NT)
    assert relmax(fg[sub], pd.synth(a)) < 10 * tol
    f = rng.standard_normal(fg.shape) + 1j * rng.standard_normal(fg.shape)
    f = np.asfortranarray(f)
    ag = pg.H @ f
    assert _rel_modes(O, ag, po.adjoint(f), L) < 1e-11
    assert np.all(ag[~O.valid_mask(L)] == 0)
    lhs = np.vdot(f, fg)
    assert abs(lhs - np.vdot(ag, a)) / abs(lhs) < 1e-12
    # the box sum runs in box order: bit-identical run to run
    assert np.array_equal(pg.H @ f, ag)


def test_three_patches(torch, P, O):
    rng = np.random.default_rng(39)
    L, tol, n = 512, 1e-8, 1500
    th = np.concatenate([rng.uniform(0.5, 0.7, n), rng.uniform(1.5, 1.7, n), rng.uniform(2.5, 2.7, n)])
    ph = np.concatenate([rng.uniform(0.5, 0.8, n), rng.uniform(3.0, 3.3, n), rng.uniform(5.0, 5.4, n)])
    pg = P.Plan(th, ph, lmax=L, tol=tol, method="multibox")
    po = O.MultiBox(th, ph, L, tol=tol, nthreads=NT)
    assert pg.info["nboxes"] == len(po.groups) == 3
    a = O.rand_alm_fast(L, rng)
    fg = pg @ a
    assert relmax(fg, po.synth(a)) < 1e-11
    f = rng.standard_normal(len(th)) + 1j * rng.standard_normal(len(th))
    ag = pg.H @ f
    assert _rel_modes(O, ag, po.adjoint(f), L) < 1e-11
    lhs = np.vdot(f, fg)
    assert abs(lhs - np.vdot(ag, a)) / abs(lhs) < 1e-12


def test_auto_takes_multibox(torch, P, O):
    # test_cost.jl:66-73: two antipodal caps at lmax 512 take the multi-box leaf; one cap does not
    rng = np.random.default_rng(74)
    n1 = 2000
    th, ph = _caps(rng, n1, 0.2)
    pg = P.Plan(th, ph, lmax=512, tol=1e-8)
    assert pg.algorithm == O.select_method(th, ph, 512, 1e-8) == "multibox"
    assert pg.info["nboxes"] == 2
    p1 = P.Plan(th[:n1], ph[:n1], lmax=512, tol=1e-8)
    assert p1.algorithm == O.select_method(th[:n1], ph[:n1], 512, 1e-8) == "stretch"


def test_analysis_on_multibox(torch, P, O):
    # the iterative callers take any plan (src/Analysis.jl): apply_resolution against the oracle
    rng = np.random.default_rng(40)
    L, tol = 128, 1e-8
    th, ph = _caps(rng, 2000, 0.4)
    pg = P.Plan(th, ph, lmax=L, tol=tol, method="multibox")
    po = O.MultiBox(th, ph, L, tol=tol, nthreads=NT)
    w = rng.uniform(0.2, 1.8, len(th))
    alm = np.asfortranarray(rng.standard_normal((L + 1, 2 * L + 1)) + 1j * rng.standard_normal((L + 1, 2 * L + 1)))
    out = pg.apply_resolution(alm, w)
    assert relmax(out, O.apply_resolution(po, alm, w)) < 1e-11
    # CG over a fixed number of iterations: the iterates agree to round-off growth
    f = pg @ O.rand_alm_fast(L, rng, cl_power=2.0)
    kw = dict(mode="cg", rtol=1e-14, maxiter=8, w=w)
    xg, ig = pg.solve(f, **kw)
    xo, io = O.solve(po, f, **kw)
    assert ig["iters"] == io["iters"] == 8
    assert _rel_modes(O, xg, xo, L) < 1e-6
