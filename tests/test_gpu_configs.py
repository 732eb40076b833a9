"""Device parity at the batched and mixed-precision workloads, through the C ABI.

- Batched: 20 deg polar cap, lmax 1024, M = 5.1e5 area-uniform points, ntrans 16 with C_l = (l+1)^-2
  spectra. All 16 columns against single applies of an ntrans = 1 plan on the same points, four of
  them against the oracle's FP64 Direct leaf on a 1e3-point subset (<= 10 tol, the bar of
  test_transforms.jl:59-82), and the adjoint identity per column (test_transforms.jl:111-123).
- Small plans: lmax 256, cap radius uniform in 5-20 deg, M = 4 (L theta*)^2, method chosen by the
  planner, against the oracle's plan of the same points (a cap below 8.25 deg is refused by the
  stretch and must go to Direct).
- Mixed precision: 30 deg polar cap, lmax 2048, M = 4 (L theta*)^2 = 4.6e6, tol 1e-5: against the
  FP64 Direct leaf on a 1e4-point subset in both directions (<= 10 tol) and against the FP64 plan
  (<= tol); mirrors gpu/runtests.jl:122-139 (a Float32 plan at tol 1e-5).
"""
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


def rel_modes(O, a, b, L):
    m = O.valid_mask(L)
    a = np.asarray(a).reshape(m.shape + (-1,), order="F")[m]
    b = np.asarray(b).reshape(m.shape + (-1,), order="F")[m]
    return float(np.abs(a - b).max() / np.abs(b).max())


def spectrum_alm(O, L, B, seed0):
    """B coefficient columns with C_l = (l+1)^-2 and distinct seeds (a_l0 real, padding zero)."""
    cols = [O.rand_alm_fast(L, np.random.default_rng(seed0 + b), cl_power=2.0) for b in range(B)]
    return np.asfortranarray(np.stack(cols, axis=2))


def test_config3_batched_lmax1024_ntrans16(torch, P, O):
    L, B, M, tol = 1024, 16, 510_000, 1e-8
    rng = np.random.default_rng(1024)
    th, ph = cap_points(rng, 20, M)
    print(f"\nbatched plan lmax {L} M {M} ntrans {B}", flush=True)
    pB = P.Plan(th, ph, lmax=L, tol=tol, ntrans=B, method="stretch")
    p1 = P.Plan(th, ph, lmax=L, tol=tol, method="stretch")
    LB = (L + 1) * (2 * L + 1)
    a = spectrum_alm(O, L, B, 5000)
    a_d = P.to_device(a)
    f_d = torch.empty(M * B, dtype=torch.complex128, device="cuda")
    pB.synth(f_d, a_d)
    f1 = torch.empty(M, dtype=torch.complex128, device="cuda")
    worst = 0.0
    for b in range(B):
        p1.synth(f1, a_d[b * LB:(b + 1) * LB])
        worst = max(worst, (f_d[b * M:(b + 1) * M] - f1).abs().max().item() / f1.abs().max().item())
    print(f"synthesis: batched vs single applies, worst column {worst:.2e}", flush=True)
    assert worst < 1e-11
    fB = P.from_device(f_d, (M, B))
    sub = np.random.default_rng(7).choice(M, 1000, replace=False)
    pd = O.Plan(th[sub], ph[sub], L, method="direct", nthreads=NT)
    for b in (0, 5, 10, 15):
        fd = pd.synth(np.asfortranarray(a[:, :, b]))
        e = relmax(fB[sub, b], fd)
        print(f"synthesis column {b} vs oracle Direct on 1e3 points: {e:.2e}", flush=True)
        assert e < 10 * tol
    # adjoint: a field supported on the subset (Direct is exact for it), then a dense field
    fs = np.random.default_rng(8).standard_normal((1000, B)) + 1j * np.random.default_rng(9).standard_normal((1000, B))
    fz = np.zeros((M, B), complex)
    fz[sub] = fs
    aB = pB.H @ np.asfortranarray(fz)
    for b in (0, 7, 15):
        e = rel_modes(O, aB[:, :, b], pd.adjoint(np.ascontiguousarray(fs[:, b])), L)
        print(f"adjoint column {b} vs oracle Direct: {e:.2e}", flush=True)
        assert e < 10 * tol
    assert np.all(aB[~O.valid_mask(L)] == 0)
    g = np.asfortranarray(np.random.default_rng(10).standard_normal((M, B))
                          + 1j * np.random.default_rng(11).standard_normal((M, B)))
    g_d = P.to_device(g)
    ag_d = torch.empty(LB * B, dtype=torch.complex128, device="cuda")
    pB.adjoint(ag_d, g_d)
    a1 = torch.empty(LB, dtype=torch.complex128, device="cuda")
    worst = 0.0
    for b in range(B):
        p1.adjoint(a1, g_d[b * M:(b + 1) * M])
        worst = max(worst, (ag_d[b * LB:(b + 1) * LB] - a1).abs().max().item() / a1.abs().max().item())
    print(f"adjoint: batched vs single applies, worst column {worst:.2e}", flush=True)
    assert worst < 1e-11
    ag = P.from_device(ag_d, (L + 1, 2 * L + 1, B))
    for b in range(B):
        lhs = np.vdot(g[:, b], fB[:, b])
        rhs = np.vdot(ag[:, :, b], a[:, :, b])
        assert abs(lhs - rhs) / abs(lhs) < 1e-12


def test_config3b_small_plans(torch, P, O):
    L, tol = 256, 1e-8
    rng = np.random.default_rng(256)
    radii = list(rng.uniform(5.0, 20.0, 6)) + [6.0, 19.5]  # at least one refused cap (< 8.25 deg)
    for i, deg in enumerate(radii):
        M = int(round(4 * (L * np.deg2rad(deg)) ** 2))
        th, ph = cap_points(rng, deg, M)
        pg = P.Plan(th, ph, lmax=L, tol=tol, method="auto")
        po = O.Plan(th, ph, L, tol=tol, method="auto", nthreads=NT)
        assert pg.algorithm == po.method, (deg, pg.algorithm, po.method)
        if deg < 8.0:
            assert pg.algorithm == "direct"
        a = O.rand_alm_fast(L, rng, cl_power=2.0)
        f = rng.standard_normal(M) + 1j * rng.standard_normal(M)
        es = relmax(pg @ a, po.synth(a))
        ea = rel_modes(O, pg.H @ f, po.adjoint(f), L)
        print(f"small plan {i}: {deg:.1f} deg, M {M}, {pg.algorithm}: synthesis {es:.1e}, adjoint {ea:.1e}",
              flush=True)
        assert es < 1e-10 and ea < 1e-10


def test_config4_mixed_cap30_lmax2048(torch, P, O):
    L, tol, deg = 2048, 1e-5, 30.0
    M = int(round(4 * (L * np.deg2rad(deg)) ** 2))
    rng = np.random.default_rng(30)
    th, ph = cap_points(rng, deg, M)
    print(f"\ncap30 lmax {L} M {M} tol {tol}", flush=True)
    pmx = P.Plan(th, ph, lmax=L, tol=tol, method="stretch", precision="mixed")
    p64 = P.Plan(th, ph, lmax=L, tol=tol, method="stretch")
    assert pmx.info["precision"] == 1 and pmx.info["kind_name"] == "pole"
    a = O.rand_alm_fast(L, rng, cl_power=2.0)
    fm, f64 = pmx @ a, p64 @ a
    e64 = relmax(fm, f64)
    sub = np.random.default_rng(31).choice(M, 10_000, replace=False)
    pdir = P.Plan(th[sub], ph[sub], lmax=L, method="direct")  # the FP64 Direct leaf on the device
    fd = pdir @ a
    ed = relmax(fm[sub], fd)
    # the device Direct leaf against the oracle's on 300 of those points
    s3 = sub[:300]
    po = O.Plan(th[s3], ph[s3], L, method="direct", nthreads=NT)
    assert relmax(fd[:300], po.synth(a)) < 1e-11
    fs = np.random.default_rng(32).standard_normal(10_000) + 1j * np.random.default_rng(33).standard_normal(10_000)
    fz = np.zeros(M, complex)
    fz[sub] = fs
    am, a64 = pmx.H @ fz, p64.H @ fz
    ad = pdir.H @ fs
    ea = rel_modes(O, am, ad, L)
    ea64 = rel_modes(O, am, a64, L)
    print(f"cap30 mixed: synthesis vs Direct {ed:.2e}, vs FP64 plan {e64:.2e}; adjoint vs Direct {ea:.2e}, "
          f"vs FP64 plan {ea64:.2e}", flush=True)
    assert ed < 10 * tol and ea < 10 * tol
    assert e64 < tol and ea64 < tol
    assert np.all(am[~O.valid_mask(L)] == 0)
