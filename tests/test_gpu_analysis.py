"""Analysis on a device plan (src/Analysis.jl) through the C ABI: apply_resolution!,
pseudo_spectrum! and solve!, against the oracle's restatements on the same inputs and against the
properties the reference's own tests assert (test_analysis.jl:122-221). Tolerances are written in
each test."""
import os

import numpy as np
import pytest

from conftest import cap_points, relmax

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


def _reldiff(a, b, O, L):
    m = O.valid_mask(L)
    return float(np.abs(a[m] - b[m]).max() / np.abs(b[m]).max())


@pytest.mark.parametrize("method", ["direct", "stretch"])
def test_apply_resolution(torch, P, O, rng, method):
    # test_analysis.jl:198-221: A' W A alm equals the adjoint of the weighted synthesis, and is
    # self-adjoint; against the oracle's plan on the same inputs
    L, n = (8, 300) if method == "direct" else (128, 6000)
    th, ph = cap_points(rng, 25, n)
    pg = P.Plan(th, ph, lmax=L, tol=1e-8, method=method)
    po = O.Plan(th, ph, L, tol=1e-8, method=method, nufft="es", nthreads=NT)
    w = rng.uniform(0.2, 1.8, n)
    alm = np.asfortranarray(rng.standard_normal((L + 1, 2 * L + 1)) + 1j * rng.standard_normal((L + 1, 2 * L + 1)))
    out = pg.apply_resolution(alm, w)
    manual = pg.H @ np.asfortranarray((pg @ alm) * w)
    assert np.array_equal(out, manual)  # the same kernels, composed on the device
    assert relmax(out, O.apply_resolution(po, alm, w)) < 1e-11
    a2 = np.asfortranarray(rng.standard_normal(alm.shape) + 1j * rng.standard_normal(alm.shape))
    o2 = pg.apply_resolution(a2, w)
    m = O.valid_mask(L)
    lhs, rhs = np.vdot(out[m], a2[m]), np.vdot(alm[m], o2[m])
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)
    # device arrays, and the default (unit) weights
    od = pg.apply_resolution(P.to_device(alm))
    assert np.array_equal(P.from_device(od, alm.shape), pg.H @ (pg @ alm))


def test_pseudo_spectrum(torch, P, O, rng):
    L, n, B = 128, 4000, 2
    th, ph = cap_points(rng, 25, n)
    pg = P.Plan(th, ph, lmax=L, tol=1e-8, ntrans=B, method="stretch")
    po = O.Plan(th, ph, L, tol=1e-8, ntrans=B, method="stretch", nufft="es", nthreads=NT)
    f = np.asfortranarray(rng.standard_normal((n, B)) + 1j * rng.standard_normal((n, B)))
    w = rng.uniform(0.0, 1.0, n)
    for b in (1, 2):
        Cg = pg.pseudo_spectrum(f, w, batch=b)
        assert relmax(Cg, O.pseudo_spectrum(po, f, w, batch=b)) < 1e-11
        a = pg.H @ np.asfortranarray(f * w[:, None])
        ab = a[:, :, b - 1]
        ref = np.array([(abs(ab[l, 0]) ** 2 + sum(abs(ab[l - m, 2 * m]) ** 2 + abs(ab[l - m, 2 * m - 1]) ** 2
                                                  for m in range(1, l + 1))) / (2 * l + 1) for l in range(L + 1)])
        assert relmax(Cg, ref) < 1e-13
    Cd = pg.pseudo_spectrum(P.to_device(f), torch.as_tensor(w, device="cuda"), batch=2)
    assert np.array_equal(Cd.cpu().numpy(), pg.pseudo_spectrum(f, w, batch=2))
    with pytest.raises(P.ArgumentError):
        pg.pseudo_spectrum(f, w, batch=3)


def test_solve_cg_null_space_and_errors(torch, P, O):
    # test_analysis.jl:122-180 on the device Direct leaf: the stopping rule is the regularisation
    rng = np.random.default_rng(52)
    L, n = 8, 400
    th = 0.6 * np.sqrt(rng.uniform(size=n))
    ph = 2 * np.pi * rng.uniform(size=n)
    plan = P.Plan(th, ph, lmax=L, method="direct")
    f = plan @ O.rand_alm_fast(L, rng)
    rec, info = plan.solve(f, mode="cg", rtol=1e-10, maxiter=2000)
    assert info["mode"] == "cg" and info["iters"] >= 1
    box = O.bounding_box(th, ph)
    assert abs(info["shannon"] - (L + 1) ** 2 * box["fsky"]) < 1e-12 * info["shannon"]
    assert abs(info["dl"] - box["dl"]) < 1e-12 * box["dl"]
    assert info["shannon"] < 0.2 * (L + 1) ** 2
    r_short, i_short = plan.solve(f, mode="cg", rtol=1e-10, maxiter=50)
    res_short = np.linalg.norm(plan @ r_short - f) / np.linalg.norm(f)
    res_long = np.linalg.norm(plan @ rec - f) / np.linalg.norm(f)
    assert i_short["iters"] == 50 and i_short["relres"] > info["relres"]
    assert res_long < res_short and res_long < 1e-4
    with pytest.raises(P.ArgumentError):
        plan.solve(f, mode="wiener")
    prior = np.array([1.0 / (1 + l) ** 2 for l in range(L + 1)])
    r2, i2 = plan.solve(f, mode="wiener", prior=prior, rtol=1e-8)
    assert i2["mode"] == "wiener" and i2["iters"] >= 1
    with pytest.raises(P.ArgumentError):
        plan.solve(f, mode="nonsense")


def test_solve_converges_on_covered_sphere(torch, P, O):
    # test_analysis.jl:152-168
    rng = np.random.default_rng(56)
    ng, L = 3000, 6
    th = np.arccos(2 * rng.uniform(size=ng) - 1)
    ph = 2 * np.pi * rng.uniform(size=ng)
    pg = P.Plan(th, ph, lmax=L, method="direct")
    at = O.rand_alm_fast(L, rng)
    fg = pg @ at
    rg, ig = pg.solve(fg, mode="cg", rtol=1e-12, maxiter=500)
    assert ig["shannon"] > 0.8 * 49
    assert np.linalg.norm(pg @ rg - fg) / np.linalg.norm(fg) < 1e-9
    assert _reldiff(rg, at, O, L) < 1e-8


@pytest.mark.parametrize("mode", ["wiener", "cg", "minnorm"])
def test_solve_vs_oracle_stretched(torch, P, O, rng, mode):
    # the same CG on the device stretched leaf and on the oracle's: the iterates agree to round-off
    # growth over a fixed number of iterations (the Wiener system is well conditioned, so its
    # solution agrees at the solver tolerance)
    L, n = 128, 4000
    th, ph = cap_points(rng, 25, n)
    pg = P.Plan(th, ph, lmax=L, tol=1e-8, method="stretch")
    po = O.Plan(th, ph, L, tol=1e-8, method="stretch", nufft="es", nthreads=NT)
    f = pg @ O.rand_alm_fast(L, rng, cl_power=2.0)
    f = f + 1e-3 * (rng.standard_normal(n) + 1j * rng.standard_normal(n))
    w = rng.uniform(0.5, 1.5, n)
    # a prior of the size of the data term (A' W A ~ M / (4 pi fsky) on the cap) keeps the Wiener
    # system well conditioned
    prior = 1e-3 / (1.0 + np.arange(L + 1) / 64.0)
    kw = dict(mode=mode, rtol=1e-8 if mode == "wiener" else 1e-14, maxiter=100 if mode == "wiener" else 12,
              prior=prior if mode == "wiener" else None, w=w)
    xg, ig = pg.solve(f, **kw)
    xo, io = O.solve(po, f, **kw)
    if mode == "wiener":
        assert ig["relres"] <= 1e-8 and abs(ig["iters"] - io["iters"]) <= 1
        assert _reldiff(xg, xo, O, L) < 1e-7
    else:
        assert ig["iters"] == io["iters"] == 12
        assert _reldiff(xg, xo, O, L) < 1e-6
        assert abs(ig["relres"] - io["relres"]) < 1e-5 * io["relres"]
    # deterministic: the same solve twice gives the same bits, on device arrays too
    xg2, ig2 = pg.solve(P.to_device(f), alm=P.to_device(np.zeros_like(xg)),
                        w=torch.as_tensor(w, device="cuda"), **{k: v for k, v in kw.items() if k != "w"})
    assert np.array_equal(P.from_device(xg2, xg.shape), xg) and ig2["iters"] == ig["iters"]
