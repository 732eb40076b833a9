"""Device parity of libpfsht_hip against the CPU oracle, through the C ABI.

Mirrors gpu/runtests.jl of the reference (ring pass forward/adjoint, end-to-end device vs host,
adjoint identity on device) and the acceptance checks: synthesis and adjoint against the oracle's
stretched leaf on the same inputs (within tol), against the FP64 Direct oracle (<= 10 tol at
small sizes, <= 1e-8 at lmax 2048, M = 2e6), the adjoint identity to round-off, and bit-identical
results run to run. Tolerances are written in each test.
"""
import os
import sys

import numpy as np
import pytest

from conftest import cap_points, centered_cap_points, relmax

pytestmark = pytest.mark.gpu

NT = min(16, os.cpu_count() or 1)


@pytest.fixture(scope="module")
def torch():
    import torch as T
    assert T.cuda.is_available(), "a GPU is required for the gpu-marked tests"
    print(f"\ndevice: {T.cuda.get_device_name(0)}", flush=True)
    return T


@pytest.fixture(scope="module")
def P():
    import pfsht
    pfsht.lib()
    return pfsht


def dev(torch, a):
    import pfsht
    return pfsht.to_device(a)


def host(t, shape):
    import pfsht
    return pfsht.from_device(t, shape)


def valid(O, L):
    return O.valid_mask(L)


def rel_modes(O, a, b, L):
    m = valid(O, L)
    a = np.asarray(a).reshape(m.shape + (-1,), order="F")[m]
    b = np.asarray(b).reshape(m.shape + (-1,), order="F")[m]
    return float(np.abs(a - b).max() / np.abs(b).max())


# --------------------------------------------------------------------------------------------
# Ring engine (gpu/runtests.jl:42-66; test_legendre.jl:87-98)
# --------------------------------------------------------------------------------------------
# (lmax, rings, transforms): one 64-ring tile; two-tile groups with several degree segments (200
# rings: tiles 64, 64, 64, 8); an odd tile count, whose last group has one tile (150 rings); two
# transforms (the FP64 MFMA passes, one column block) and five (two column blocks, four + one)
@pytest.mark.parametrize("L,nr,B", [(128, 64, 1), (512, 64, 1), (300, 200, 1), (300, 150, 1), (300, 200, 2),
                                    (260, 150, 5)])
def test_ring_forward_and_adjoint_vs_oracle(torch, P, O, rng, L, nr, B):
    th = np.sort(0.02 + 0.5 * rng.uniform(size=nr))
    a = O.rand_alm_fast(L, rng, ntrans=B) if B > 1 else O.rand_alm_fast(L, rng)
    Gh = O.Rings(th, L, 1e-8, ntrans=B).rings_from_alm(a)
    eng = P.Rings(th, L, 1e-8, ntrans=B)
    assert eng.work == O.Rings(th, L, 1e-8).work and eng.mmax == O.Rings(th, L, 1e-8).mmax
    Gd = torch.zeros(Gh.size, dtype=torch.complex128, device="cuda")
    eng.rings_from_alm(Gd, dev(torch, a))
    torch.cuda.synchronize()
    G = host(Gd, Gh.shape)
    # per-ring exponents and FMA contraction versus the host's shared chunk exponent: round-off
    assert relmax(G, Gh) < 1e-12
    assert np.all(G[Gh == 0] == 0)
    Gr = rng.standard_normal(Gh.shape) + 1j * rng.standard_normal(Gh.shape)
    Gr[Gh == 0] = 0  # entries off the retained set are never read
    ah = O.Rings(th, L, 1e-8, ntrans=B).alm_from_rings(Gr)
    ad = torch.zeros(ah.size, dtype=torch.complex128, device="cuda")
    eng.alm_from_rings(ad, dev(torch, Gr))
    assert relmax(host(ad, ah.shape), ah) <= 1e-12


def test_ring_engine_matches_alf(torch, P, O):
    L = 48
    th = np.sort([0.02, 0.11, 0.37, 0.8, 1.4, 2.7])
    a = np.zeros((L + 1, 2 * L + 1), complex, order="F")
    a[P.sph_mode(30, 7)] = 1.0
    G = np.zeros((len(th), 2 * L + 1, 1), complex, order="F")
    P.Rings(th, L, 1e-13).rings_from_alm(G, a)
    for j, t in enumerate(th):
        assert abs(G[j, 7 + L, 0] - O.alf(30, 7, t)) <= 1e-13


def test_ring_engine_reflect_and_equator(torch, P, O, rng):
    L = 256
    for th, refl in ((np.linspace(0.5, 2.6, 77), False), (np.sort(rng.uniform(0.1, 0.9, 40)), True)):
        a = O.rand_alm_fast(L, rng)
        Gh = O.Rings(th, L, 1e-8, reflect=refl).rings_from_alm(a)
        G = np.zeros(Gh.shape, complex, order="F")
        P.Rings(th, L, 1e-8, reflect=refl).rings_from_alm(G, a)
        assert relmax(G, Gh) < 1e-12


# --------------------------------------------------------------------------------------------
# Stretched plan end to end
# --------------------------------------------------------------------------------------------
def _plans(P, O, th, ph, L, tol, B=1):
    pg = P.Plan(th, ph, lmax=L, tol=tol, ntrans=B, method="stretch")
    po = O.Plan(th, ph, L, tol=tol, ntrans=B, method="stretch", nufft="es", nthreads=NT)
    for k in ("kind", "flip", "nth_half", "nth", "nph", "mmax", "ring_work", "es_w", "es_n1", "es_n2"):
        assert pg.info[k] == po.info[k], k
    return pg, po


def test_nufft_seam_vs_oracle_and_direct_sum(torch, P, O, rng):
    # _nufft_exec! on the plan's stretched points (src/NUFFT.jl:17-25,85-129): type 2 (iflag +1)
    # and type 1 (iflag -1) against the oracle's ES NUFFT (same kernel, <1e-12) and against the
    # reference's direct sum (200 tol, test_backends.jl:1-19); and the pair is an exact adjoint.
    L, tol, M = 160, 1e-8, 700
    th, ph = cap_points(rng, 20, M)
    pg, po = _plans(P, O, th, ph, L, tol)
    T, p = po.points()
    n1, n2 = pg.info["nth"], pg.info["nph"]
    c = np.asfortranarray(rng.standard_normal((n1, n2)) + 1j * rng.standard_normal((n1, n2)))
    f = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    out2 = np.zeros(M, complex)
    pg.nufft(2, out2, c)
    assert relmax(out2, O.nufft_es(2, n1, n2, +1, T, p, c, tol)[:, 0]) < 1e-12
    assert relmax(out2, O.nufft_direct(2, n1, n2, +1, T, p, c)[:, 0]) < 200 * tol
    out1 = np.zeros((n1, n2), complex, order="F")
    pg.nufft(1, out1, f)
    assert relmax(out1, O.nufft_es(1, n1, n2, -1, T, p, f, tol)[:, :, 0]) < 1e-12
    assert relmax(out1, O.nufft_direct(1, n1, n2, -1, T, p, f)[:, :, 0]) < 200 * tol
    assert abs(np.vdot(f, out2) - np.vdot(out1, c)) < 1e-13 * abs(np.vdot(f, out2))


def test_config1_synth_adjoint_vs_oracle(torch, P, O, rng):
    # 20 deg cap, lmax 256, M = 4 (L theta*)^2, tol 1e-8, B = 1
    L, tol = 256, 1e-8
    M = int(round(4 * (L * np.deg2rad(20)) ** 2))
    th, ph = cap_points(rng, 20, M)
    pg, po = _plans(P, O, th, ph, L, tol)
    a = O.rand_alm_fast(L, rng)
    fd = torch.zeros(M, dtype=torch.complex128, device="cuda")
    ad = dev(torch, a)
    pg.synth(fd, ad)
    fg = host(fd, (M,))
    fo = po.synth(a)
    # same algorithm and kernel parameters as the oracle: agreement far inside tol
    assert relmax(fg, fo) < 1e-11
    # against the FP64 Direct oracle on a subset of points (test_transforms.jl:59-82: 10 tol)
    sub = rng.choice(M, 300, replace=False)
    pd = O.Plan(th[sub], ph[sub], L, method="direct", nthreads=NT)
    assert relmax(fg[sub], pd.synth(a)) < 10 * tol
    # adjoint
    f = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    alm_d = torch.zeros(a.size, dtype=torch.complex128, device="cuda")
    pg.adjoint(alm_d, dev(torch, f))
    ag = host(alm_d, a.shape)
    ao = po.adjoint(f)
    assert rel_modes(O, ag, ao, L) < 1e-11
    assert np.all(ag[~O.valid_mask(L)] == 0)  # padding overwritten with zeros
    # adjoint identity <Y a, f> = <a, Y' f> to round-off (target ~1e-13; the reference's gate is
    # 1e-12, test/test_transforms.jl:111-123)
    lhs = np.vdot(f, fg)
    rhs = np.vdot(ag, a)
    assert abs(lhs - rhs) / abs(lhs) < 1e-12


def test_bit_identical_run_to_run(torch, P, O, rng):
    L, tol = 256, 1e-8
    th, ph = cap_points(rng, 20, 20000)
    pg = P.Plan(th, ph, lmax=L, tol=tol)
    a = dev(torch, O.rand_alm_fast(L, rng))
    f = dev(torch, rng.standard_normal(20000) + 1j * rng.standard_normal(20000))
    outs_f, outs_a = [], []
    for _ in range(3):
        fo = torch.empty(20000, dtype=torch.complex128, device="cuda")
        ao = torch.empty(a.numel(), dtype=torch.complex128, device="cuda")
        pg.synth(fo, a)
        pg.adjoint(ao, f)
        outs_f.append(fo)
        outs_a.append(ao)
    torch.cuda.synchronize()
    for k in (1, 2):
        assert torch.equal(outs_f[0], outs_f[k])
        assert torch.equal(outs_a[0], outs_a[k])


@pytest.mark.parametrize("L,tc,kind,flip", [(256, 0.9, "pole", 0), (256, np.pi - 0.3, "pole", 1),
                                             (512, 1.2, "interior", 0)])
def test_off_pole_south_interior(torch, P, O, rng, L, tc, kind, flip):
    # test_transforms.jl:125-140
    tol, n = 1e-8, 150
    th = tc + 0.12 * (2 * rng.uniform(size=n) - 1)
    ph = 1.0 + 0.4 * rng.uniform(size=n)
    pg, po = _plans(P, O, th, ph, L, tol)
    assert pg.info["kind_name"] == kind and pg.info["flip"] == flip
    a = O.rand_alm_fast(L, rng)
    fg = pg @ a
    pd = O.Plan(th, ph, L, method="direct")
    assert relmax(fg, pd.synth(a)) < 10 * tol
    assert relmax(fg, po.synth(a)) < 1e-11
    f = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    ag = pg.H @ f
    assert abs(np.vdot(f, fg) - np.vdot(ag, a)) / abs(np.vdot(f, fg)) < 1e-12
    assert rel_modes(O, ag, po.adjoint(f), L) < 1e-11


@pytest.mark.parametrize("deg,tc,L", [(15, np.deg2rad(60), 512), (30, np.pi / 2, 256)])
def test_interior_caps_vs_direct(torch, P, O, rng, deg, tc, L):
    # off-pole caps whose colatitude range stays clear of both poles: the interior periodisation
    # (src/Windows.jl:84-141), at M = 4 (L theta*)^2; synthesis against the FP64 Direct oracle on a
    # subset (10 tol, test_transforms.jl:59-82), the adjoint against Direct for a field supported on
    # that subset, and the adjoint identity.
    tol = 1e-8
    M = int(round(4 * (L * np.deg2rad(deg)) ** 2))
    th, ph = centered_cap_points(rng, deg, M, tc, 1.0)
    pg = P.Plan(th, ph, lmax=L, tol=tol, method="stretch")
    assert pg.info["kind_name"] == "interior"
    a = O.rand_alm_fast(L, rng, cl_power=2.0)
    fg = pg @ a
    sub = rng.choice(M, 300, replace=False)
    pd = O.Plan(th[sub], ph[sub], L, method="direct", nthreads=NT)
    assert relmax(fg[sub], pd.synth(a)) < 10 * tol
    fs = rng.standard_normal(300) + 1j * rng.standard_normal(300)
    f = np.zeros(M, complex)
    f[sub] = fs
    ag = pg.H @ f
    assert rel_modes(O, ag, pd.adjoint(fs), L) < 10 * tol
    f = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    ag = pg.H @ f
    lhs = np.vdot(f, fg)
    assert abs(lhs - np.vdot(ag, a)) / abs(lhs) < 1e-12


def test_ntrans_columns(torch, P, O, rng):
    # test_transforms.jl:142-156
    L, B, n = 256, 3, 4000
    th, ph = cap_points(rng, 20, n)
    p1 = P.Plan(th, ph, lmax=L, tol=1e-8)
    pB = P.Plan(th, ph, lmax=L, tol=1e-8, ntrans=B)
    aB = O.rand_alm_fast(L, rng, ntrans=B)
    fB = pB @ aB
    fs = np.stack([p1 @ np.asfortranarray(aB[:, :, b]) for b in range(B)], axis=1)
    assert relmax(fB, fs) < 1e-13
    f = rng.standard_normal((n, B)) + 1j * rng.standard_normal((n, B))
    f = np.asfortranarray(f)
    aBg = pB.H @ f
    for b in range(B):
        assert rel_modes(O, aBg[:, :, b], p1.H @ np.ascontiguousarray(f[:, b]), L) < 1e-13


def test_ntrans16_columns(torch, P, O, rng):
    # the batched form (two groups of 8 transforms per ring launch) equals sixteen single applies
    L, B, n = 128, 16, 2500
    th, ph = cap_points(rng, 20, n)
    p1 = P.Plan(th, ph, lmax=L, tol=1e-8, method="stretch")
    pB = P.Plan(th, ph, lmax=L, tol=1e-8, ntrans=B, method="stretch")
    aB = O.rand_alm_fast(L, rng, ntrans=B)
    fB = pB @ aB
    for b in range(B):
        assert relmax(fB[:, b], p1 @ np.asfortranarray(aB[:, :, b])) < 1e-13
    f = np.asfortranarray(rng.standard_normal((n, B)) + 1j * rng.standard_normal((n, B)))
    aBg = pB.H @ f
    for b in range(B):
        assert rel_modes(O, aBg[:, :, b], p1.H @ np.ascontiguousarray(f[:, b]), L) < 1e-13
    assert np.all(aBg[~O.valid_mask(L)] == 0)


@pytest.mark.parametrize("tol,deg,tc,ntrans", [(1e-6, 20, 0.0, 1), (1e-5, 25, 1.1, 2)])
def test_mixed_precision(torch, P, O, rng, tol, deg, tc, ntrans):
    # mixed precision (tol >= 1e-6): FP64 ring recurrence, grid buffers and FFTs in FP32. Against
    # the FP64 plan at the same tol (within tol), the FP64 Direct oracle on a subset (10 tol, the
    # bar of test_transforms.jl:59-82), and the adjoint identity (within tol: FP32 round-off).
    L = 256
    M = int(round(4 * (L * np.deg2rad(deg)) ** 2))
    th, ph = centered_cap_points(rng, deg, M, tc, 0.5) if tc else cap_points(rng, deg, M)
    p64 = P.Plan(th, ph, lmax=L, tol=tol, ntrans=ntrans, method="stretch")
    pmx = P.Plan(th, ph, lmax=L, tol=tol, ntrans=ntrans, method="stretch", precision="mixed")
    assert pmx.info["precision"] == 1 and pmx.info["kind"] == p64.info["kind"]
    a = O.rand_alm_fast(L, rng, ntrans=ntrans)
    fm = pmx @ a
    f64 = p64 @ a
    e64 = relmax(fm, f64)
    sub = rng.choice(M, 200, replace=False)
    pd = O.Plan(th[sub], ph[sub], L, method="direct", nthreads=NT)
    cols = [np.asfortranarray(a[:, :, b]) for b in range(ntrans)] if ntrans > 1 else [a]
    fd = np.stack([pd.synth(c) for c in cols], axis=1).reshape(fm[sub].shape)
    ed = relmax(fm[sub], fd)
    fs = rng.standard_normal((200, ntrans)) + 1j * rng.standard_normal((200, ntrans))
    f = np.zeros((M, ntrans), complex)
    f[sub] = fs
    f = np.asfortranarray(f if ntrans > 1 else f[:, 0])
    am = pmx.H @ f
    ea = max(rel_modes(O, (am[:, :, b] if ntrans > 1 else am),
                       pd.adjoint(np.ascontiguousarray(fs[:, b])), L) for b in range(ntrans))
    g = np.asfortranarray(rng.standard_normal(f.shape) + 1j * rng.standard_normal(f.shape))
    ag = pmx.H @ g
    lhs = np.vdot(g, fm)
    eid = abs(lhs - np.vdot(ag, a)) / abs(lhs)
    print(f"mixed tol={tol}: vs fp64 plan {e64:.2e}, vs Direct {ed:.2e}, adjoint vs Direct {ea:.2e}, "
          f"identity {eid:.2e}", flush=True)
    assert e64 < tol
    assert ed < 10 * tol
    assert ea < 10 * tol
    assert eid < tol
    assert np.all(am[~O.valid_mask(L)] == 0)
    with pytest.raises(P.ArgumentError):   # the NUFFT seam is FP64 only
        pmx.nufft(2, np.zeros(M * ntrans, complex), np.zeros(pmx.info["nth"] * pmx.info["nph"] * ntrans, complex))
    with pytest.raises(P.ArgumentError):
        P.Plan(th, ph, lmax=L, tol=1e-8, method="stretch", precision="mixed")


def test_shape_errors(torch, P, rng):
    th, ph = cap_points(rng, 20, 1000)
    p = P.Plan(th, ph, lmax=160, tol=1e-8, method="stretch")
    with pytest.raises(P.DimensionMismatch):
        p.synth(np.zeros(999, complex), P.zeros_alm(160))
    with pytest.raises(P.DimensionMismatch):
        p.synth(np.zeros(1000, complex), P.zeros_alm(159))
    with pytest.raises(P.DimensionMismatch):
        p.adjoint(P.zeros_alm(160, ntrans=2), np.zeros(1000, complex))


def test_refused_geometry(torch, P, rng):
    th, ph = cap_points(rng, 5, 2000)  # 5 deg cap at lmax 256, tol 1e-8: no admissible stretch
    with pytest.raises(P.ArgumentError):
        P.Plan(th, ph, lmax=256, tol=1e-8, method="stretch")


# --------------------------------------------------------------------------------------------
# Direct leaf on the device (src/Direct.jl; test_transforms.jl:1-57)
# --------------------------------------------------------------------------------------------
def test_direct_leaf_vs_oracle(torch, P, O, rng):
    # whole-sphere points with both poles, the equator and repeated colatitudes, B = 2
    L, B, n = 128, 2, 600
    th = np.arccos(2 * rng.uniform(size=n) - 1)
    th[:4] = [0.0, np.pi, np.pi / 2, np.pi / 2]
    th[4:8] = th[8]
    ph = 2 * np.pi * rng.uniform(size=n)
    pg = P.Plan(th, ph, lmax=L, ntrans=B, method="direct")
    po = O.Plan(th, ph, L, ntrans=B, method="direct", nthreads=NT)
    assert pg.algorithm == "direct" and pg.info["kind_name"] == "direct"
    assert pg.info["mmax"] == L
    a = O.rand_alm_fast(L, rng, ntrans=B)
    fg = pg @ a
    fo = po.synth(a)
    assert relmax(fg, fo) < 1e-12
    f = np.asfortranarray(rng.standard_normal((n, B)) + 1j * rng.standard_normal((n, B)))
    ag = pg.H @ f
    ao = po.adjoint(f)
    for b in range(B):
        assert rel_modes(O, ag[:, :, b], ao[:, :, b], L) < 1e-12
        lhs = np.vdot(f[:, b], fg[:, b])
        assert abs(lhs - np.vdot(ag[:, :, b], a[:, :, b])) / abs(lhs) < 1e-12
    assert np.all(ag[~O.valid_mask(L)] == 0)


def test_auto_selection_matches_oracle(torch, P, O, rng):
    # src/Plan.jl:84-116 with src/Cost.jl:174-201: a tiny set goes Direct, a cap the stretch
    # refuses goes Direct, a large cap goes Stretch; the device plan agrees with the oracle's
    # planner and with its Direct results.
    cases = [(64, 2, 20), (256, 5, 2000), (256, 20, 30000)]
    for L, deg, n in cases:
        th, ph = cap_points(rng, deg, n)
        pg = P.Plan(th, ph, lmax=L, tol=1e-8)
        po = O.Plan(th, ph, L, tol=1e-8, nthreads=NT)
        assert pg.algorithm == po.method, (L, deg, n)
    th, ph = cap_points(rng, 5, 2000)
    pg = P.Plan(th, ph, lmax=256, tol=1e-8)
    assert pg.algorithm == "direct"
    a = O.rand_alm_fast(256, rng)
    pd = O.Plan(th, ph, 256, method="direct", nthreads=NT)
    assert relmax(pg @ a, pd.synth(a)) < 1e-12


# --------------------------------------------------------------------------------------------
# Serving runtime: plan cache, admission, applies over several streams
# --------------------------------------------------------------------------------------------
# --------------------------------------------------------------------------------------------
# Full size: lmax 2048, M = 2e6 (the benchmark workload)
# --------------------------------------------------------------------------------------------
def test_full_size_config2(torch, P, O, rng):
    L, tol, M = 2048, 1e-8, 2_000_000
    th, ph = cap_points(rng, 20, M)
    print("building lmax 2048 plan", flush=True)
    pg = P.Plan(th, ph, lmax=L, tol=tol, method="stretch")
    a = O.rand_alm_fast(L, rng, cl_power=2.0)
    ad = dev(torch, a)
    fd = torch.zeros(M, dtype=torch.complex128, device="cuda")
    pg.synth(fd, ad)
    fg = host(fd, (M,))
    sub = rng.choice(M, 256, replace=False)
    print("direct oracle on a 256-point subset", flush=True)
    pd = O.Plan(th[sub], ph[sub], L, method="direct", nthreads=NT)
    fdir = pd.synth(a)
    err_s = relmax(fg[sub], fdir)
    print(f"synthesis vs Direct: {err_s:.3e}", flush=True)
    assert err_s <= 1e-8
    # adjoint with f supported on the subset: Y'f then involves only those points
    f = np.zeros(M, complex)
    f[sub] = rng.standard_normal(256) + 1j * rng.standard_normal(256)
    alm_d = torch.zeros(a.size, dtype=torch.complex128, device="cuda")
    pg.adjoint(alm_d, dev(torch, f))
    ag = host(alm_d, a.shape)
    adir = pd.adjoint(f[sub])
    err_a = rel_modes(O, ag, adir, L)
    print(f"adjoint vs Direct: {err_a:.3e}", flush=True)
    assert err_a <= 1e-8
    # adjoint identity on a dense random field
    g = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    pg.adjoint(alm_d, dev(torch, g))
    ag = host(alm_d, a.shape)
    lhs, rhs = np.vdot(g, fg), np.vdot(ag, a)
    rid = abs(lhs - rhs) / abs(lhs)
    print(f"adjoint identity: {rid:.3e}", flush=True)
    assert rid < 1e-12
