"""Column applies (pfsht_synth_cols / pfsht_adjoint_cols) and the plan keywords, through the C ABI.

A column apply of nb <= ntrans transforms whose columns are separate device buffers must give, for
each column, the apply of that column alone (to round-off; the single-transform and the MFMA ring
passes differ in summation order only), on every leaf kind. The keywords oversample_θ / oversample_φ
/ εapo_max / maxboxes (src/Plan.jl:84-135) change the plan as the reference's sizing says and keep it
within tol of the Direct oracle.
"""
import os

import numpy as np
import pytest

from conftest import cap_points, centered_cap_points, relmax

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


def alm_col(torch, O, L, seed, spin=0):
    if spin:
        a = np.zeros((L + 1, 2 * L + 1), complex, order="F")
        m = O.spin_valid_mask(L, spin)
        r = np.random.default_rng(seed)
        a[m] = r.standard_normal(m.sum()) + 1j * r.standard_normal(m.sum())
    else:
        a = O.rand_alm_fast(L, np.random.default_rng(seed), cl_power=2.0)
    return torch.from_numpy(np.asarray(a).ravel(order="F").copy()).cuda()


def field_col(torch, M, seed):
    r = np.random.default_rng(seed)
    return torch.from_numpy(r.standard_normal(M) + 1j * r.standard_normal(M)).cuda()


def rel(torch, a, b):
    return (a - b).abs().max().item() / b.abs().max().item()


def check_cols(torch, P, O, pB, p1, L, M, nbs, spin=0, bound=1e-11):
    """Scattered columns through the column apply against single applies of p1 (ntrans 1)."""
    LB = (L + 1) * (2 * L + 1)
    for nb in nbs:
        alms = [alm_col(torch, O, L, 100 + b, spin) for b in range(nb)]
        fs = [torch.empty(M, dtype=torch.complex128, device="cuda") for _ in range(nb)]
        pB.synth_cols(fs, alms)
        one = torch.empty(M, dtype=torch.complex128, device="cuda")
        for b in range(nb):
            p1.synth(one, alms[b])
            assert rel(torch, fs[b], one) < bound, (nb, b)
        gs = [field_col(torch, M, 200 + b) for b in range(nb)]
        outs = [torch.empty(LB, dtype=torch.complex128, device="cuda") for _ in range(nb)]
        pB.adjoint_cols(outs, gs)
        a1 = torch.empty(LB, dtype=torch.complex128, device="cuda")
        for b in range(nb):
            p1.adjoint(a1, gs[b])
            assert rel(torch, outs[b], a1) < bound, (nb, b)
    torch.cuda.synchronize()


@pytest.mark.parametrize("packed", [False, True])
def test_cols_stretch(torch, P, O, rng, packed):
    L, M, B = 256, 6000, 6
    th, ph = cap_points(rng, 20, M)
    pB = P.Plan(th, ph, lmax=L, tol=1e-8, ntrans=B, method="stretch", packed=packed)
    p1 = P.Plan(th, ph, lmax=L, tol=1e-8, method="stretch")
    check_cols(torch, P, O, pB, p1, L, M, [1, 2, 3, 5, 6])


def test_cols_direct_multibox_spin(torch, P, O, rng):
    L = 128
    th, ph = cap_points(rng, 60, 1500)
    check_cols(torch, P, O, P.Plan(th, ph, lmax=L, ntrans=4, method="direct"),
               P.Plan(th, ph, lmax=L, method="direct"), L, 1500, [1, 3, 4])
    t1, p1_ = cap_points(rng, 15, 2500)
    t2, p2_ = centered_cap_points(rng, 15, 2500, np.pi, 1.0)
    th, ph = np.concatenate([t1, t2]), np.concatenate([p1_, p2_])
    L = 256  # a 15 deg cap is below the stretch's reach at lmax 128 (L h < 2 beta)
    pm = P.Plan(th, ph, lmax=L, tol=1e-8, ntrans=3, method="multibox", packed=True)
    assert pm.algorithm == "multibox" and pm.info["nboxes"] == 2
    # the same plan's contiguous apply of the columns (zero past nb) is the reference: the partition
    # may depend on ntrans through the cost model, so a plan of another ntrans is not
    M, LB = 5000, (L + 1) * (2 * L + 1)
    for nb in (2, 3):
        alms = [alm_col(torch, O, L, 300 + b) for b in range(nb)]
        fs = [torch.empty(M, dtype=torch.complex128, device="cuda") for _ in range(nb)]
        pm.synth_cols(fs, alms)
        full = torch.cat(alms + [torch.zeros(LB, dtype=torch.complex128, device="cuda")] * (3 - nb))
        fc = torch.empty(3 * M, dtype=torch.complex128, device="cuda")
        pm.synth(fc, full)
        for b in range(nb):
            assert rel(torch, fs[b], fc[b * M:(b + 1) * M]) < 1e-12
        gs = [field_col(torch, M, 400 + b) for b in range(nb)]
        outs = [torch.empty(LB, dtype=torch.complex128, device="cuda") for _ in range(nb)]
        pm.adjoint_cols(outs, gs)
        gfull = torch.cat(gs + [torch.zeros(M, dtype=torch.complex128, device="cuda")] * (3 - nb))
        ac = torch.empty(3 * LB, dtype=torch.complex128, device="cuda")
        pm.adjoint(ac, gfull)
        for b in range(nb):
            assert rel(torch, outs[b], ac[b * LB:(b + 1) * LB]) < 1e-12
    th, ph = cap_points(rng, 20, 3000)
    L = 128
    check_cols(torch, P, O, P.Plan(th, ph, lmax=L, tol=1e-8, ntrans=3, spin=2, method="stretch"),
               P.Plan(th, ph, lmax=L, tol=1e-8, spin=2, method="stretch"), L, 3000, [1, 2, 3], spin=2)


def test_cols_mixed_precision(torch, P, O, rng):
    L, M = 256, 5000
    th, ph = cap_points(rng, 20, M)
    pB = P.Plan(th, ph, lmax=L, tol=1e-5, ntrans=4, method="stretch", precision="mixed", packed=True)
    p1 = P.Plan(th, ph, lmax=L, tol=1e-5, method="stretch", precision="mixed")
    check_cols(torch, P, O, pB, p1, L, M, [1, 3, 4], bound=1e-5)


def test_cols_arguments(torch, P, rng):
    th, ph = cap_points(rng, 20, 2000)
    p = P.Plan(th, ph, lmax=256, tol=1e-8, ntrans=2, method="stretch")
    a = torch.zeros(257 * 513, dtype=torch.complex128, device="cuda")
    f = torch.zeros(2000, dtype=torch.complex128, device="cuda")
    with pytest.raises(P.ArgumentError):  # more columns than the plan's ntrans
        p.synth_cols([f, f, f], [a, a, a])
    with pytest.raises(P.DimensionMismatch):
        p.synth_cols([torch.zeros(1999, dtype=torch.complex128, device="cuda")], [a])
    with pytest.raises(P.ArgumentError):
        p.synth_cols([f.cpu()], [a.cpu()])


def test_plan_keywords(torch, P, O, rng):
    L, tol, M = 256, 1e-8, 5000
    th, ph = cap_points(rng, 20, M)
    base = P.Plan(th, ph, lmax=L, tol=tol, method="stretch")
    over = P.Plan(th, ph, lmax=L, tol=tol, method="stretch", oversample_theta=1.5, oversample_phi=1.25)
    # src/Stretch.jl:164 and :194: ceil((K + margin) * os) before the smooth rounding
    g = O.stretch_geometry(th.min(), th.max(), L, tol)
    Kt = int(np.ceil((g.K + 2) * 1.5))
    assert over.info["nth_half"] == O.next_smooth(Kt + 1)
    assert over.info["nph"] == O.next_smooth_even(int(np.ceil((2 * base.info["mmax"] + 1) * 1.25)))
    assert over.info["nth"] > base.info["nth"] and over.info["nph"] > base.info["nph"]
    a = O.rand_alm_fast(L, rng, cl_power=2.0)
    sub = rng.choice(M, 300, replace=False)
    pd = O.Plan(th[sub], ph[sub], L, method="direct", nthreads=NT)
    assert relmax((over @ a)[sub], pd.synth(a)) < 10 * tol
    # a longitude grid past 2048 (oversample_phi 12): the fine length is past 4096 and takes the next
    # single-kernel rocFFT length; both directions stay within 10 tol of Direct
    wide = P.Plan(th, ph, lmax=L, tol=tol, method="stretch", oversample_phi=12.0)
    assert wide.info["nph"] > 2048 and wide.info["es_n2"] == O.es_fine_size(wide.info["nph"], wide.info["es_w"])
    assert wide.info["es_n2"] in (6144, 8192, 9216, 10000, 10240)
    assert relmax((wide @ a)[sub], pd.synth(a)) < 10 * tol
    fw = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    fz = np.zeros(M, complex)
    fz[sub] = fw[sub]
    m = O.valid_mask(L)
    assert relmax((wide.H @ fz)[m], pd.adjoint(fw[sub])[m]) < 10 * tol
    # εapo_max: a cap the default admits is refused below its apodisation fraction
    ea = g.epsapo
    assert O.stretch_geometry(th.min(), th.max(), L, tol, epsapo_max=0.9 * ea) is None
    with pytest.raises(P.ArgumentError):
        P.Plan(th, ph, lmax=L, tol=tol, method="stretch", epsapo_max=0.9 * ea)
    assert P.Plan(th, ph, lmax=L, tol=tol, method="stretch", epsapo_max=1.1 * ea).info["kind_name"] == "pole"
    # maxboxes = 1: two antipodal caps stay one box
    t2, p2 = centered_cap_points(rng, 15, 3000, np.pi, 1.0)
    tt, pp = np.concatenate([th[:3000], t2]), np.concatenate([ph[:3000], p2])
    assert P.Plan(tt, pp, lmax=L, tol=tol, method="auto").algorithm == "multibox"
    assert P.Plan(tt, pp, lmax=L, tol=tol, method="auto", maxboxes=1).algorithm != "multibox"
    for bad in (dict(oversample_theta=0.5), dict(epsapo_max=1.5), dict(maxboxes=-1)):
        with pytest.raises(P.ArgumentError):
            P.Plan(th, ph, lmax=L, tol=tol, **bad)
