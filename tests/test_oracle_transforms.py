"""Pins the oracle's leaves end to end (after src/Direct.jl, src/Stretch.jl, src/DFS.jl, src/NUFFT.jl).

Restates test/test_transforms.jl and test/test_backends.jl:1-19,187-200. Direct synthesis is
checked against scipy's spherical harmonics summed term by term (independent of the ring
engine); the stretched leaf against Direct at 10*tol; the ES NUFFT against the reference's own
direct-sum NUFFT at 200*tol; the FFT against numpy's.
"""
import numpy as np
import pytest
from scipy.special import sph_harm_y

from conftest import cap_points, relmax


def reladj(p, a, f):
    lhs = np.vdot(f, p.synth(a))
    rhs = np.vdot(p.adjoint(f), a)
    return abs(lhs - rhs) / abs(lhs)


def test_direct_vs_explicit_sum(O, rng):
    # test_transforms.jl:9-31
    for L in (6, 32):
        n = 40
        th = 0.4 * rng.uniform(size=n)
        ph = 2 * np.pi * rng.uniform(size=n)
        p = O.Plan(th, ph, L, method="direct")
        a = rng.standard_normal((L + 1, 2 * L + 1)) + 1j * rng.standard_normal((L + 1, 2 * L + 1))
        a = np.asfortranarray(a)
        ref = np.zeros(n, complex)
        for l in range(L + 1):
            for m in range(-l, l + 1):
                r, c = O.sph_mode(l, m)
                ref += a[r, c] * sph_harm_y(l, m, th, ph)
        f = p.synth(a)
        assert relmax(f, ref) < 1e-13
        assert reladj(p, a, rng.standard_normal(n) + 1j * rng.standard_normal(n)) < 1e-13


def test_single_harmonic(O, rng):
    # test_transforms.jl:33-43
    L, n = 8, 25
    th = 0.5 * rng.uniform(size=n) + 0.1
    ph = 2 * np.pi * rng.uniform(size=n)
    p = O.Plan(th, ph, L, method="direct")
    for (l, m) in ((0, 0), (1, -1), (5, 3), (8, -8)):
        a = np.zeros((L + 1, 2 * L + 1), complex, order="F")
        r, c = O.sph_mode(l, m)
        a[r, c] = 1
        assert np.abs(p.synth(a) - sph_harm_y(l, m, th, ph)).max() < 1e-13


def test_direct_ntrans(O, rng):
    L, n, B = 12, 30, 3
    th = 0.4 * rng.uniform(size=n)
    ph = 2 * np.pi * rng.uniform(size=n)
    p1 = O.Plan(th, ph, L, method="direct")
    pB = O.Plan(th, ph, L, ntrans=B, method="direct")
    aB = O.rand_alm_fast(L, rng, ntrans=B)
    fB = pB.synth(aB)
    for b in range(B):
        assert np.abs(fB[:, b] - p1.synth(aB[:, :, b])).max() < 1e-13


@pytest.mark.parametrize("L", [128, 256])
@pytest.mark.parametrize("tol", [1e-4, 1e-8])
def test_stretch_vs_direct(O, rng, L, tol):
    # test_transforms.jl:59-82, with the reference's exact direct-sum NUFFT, and with the ES NUFFT
    n = 150
    th = np.deg2rad(20) * np.sqrt(rng.uniform(size=n))
    ph = 2 * np.pi * rng.uniform(size=n)
    pd = O.Plan(th, ph, L, method="direct")
    a = rng.standard_normal((L + 1, 2 * L + 1)) + 1j * rng.standard_normal((L + 1, 2 * L + 1))
    a = np.asfortranarray(a)
    fd = pd.synth(a)
    for nufft in ("direct", "es"):
        ps = O.Plan(th, ph, L, tol=tol, method="stretch", nufft=nufft)
        assert relmax(ps.synth(a), fd) < 10 * tol
        assert ps.info["nph"] % 2 == 0


def test_stretch_worst_case_spectra(O, rng):
    # test_transforms.jl:84-109
    L, tol, n = 256, 1e-8, 200
    th = np.deg2rad(20) * np.sqrt(rng.uniform(size=n))
    ph = 2 * np.pi * rng.uniform(size=n)
    pd = O.Plan(th, ph, L, method="direct")
    ps = O.Plan(th, ph, L, tol=tol, method="stretch", nufft="direct")
    a = np.zeros((L + 1, 2 * L + 1), complex, order="F")
    for m in range(-L, L + 1):
        a[O.sph_mode(L, m)] = rng.standard_normal() + 1j * rng.standard_normal()
    assert relmax(ps.synth(a), pd.synth(a)) < 10 * tol
    ar = O.rand_alm_fast(L, rng, cl_power=4.0)
    assert relmax(ps.synth(ar), pd.synth(ar)) < 10 * tol


@pytest.mark.parametrize("L", [256, 512])
def test_stretch_adjoint_identity(O, rng, L):
    # test_transforms.jl:111-123: round-off, not tolerance
    n = 120
    th = np.deg2rad(20) * np.sqrt(rng.uniform(size=n))
    ph = 2 * np.pi * rng.uniform(size=n)
    a = O.rand_alm_fast(L, rng)
    f = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    for nufft in ("direct", "es"):
        ps = O.Plan(th, ph, L, tol=1e-8, method="stretch", nufft=nufft)
        assert reladj(ps, a, f) < 1e-12


@pytest.mark.parametrize("L,tc,kind,flip", [(256, 0.9, "pole", 0), (256, np.pi - 0.3, "pole", 1),
                                             (512, 1.2, "interior", 0)])
def test_stretch_off_pole_and_south(O, rng, L, tc, kind, flip):
    # test_transforms.jl:125-140 (mid-latitude and southern patches), plus a patch where the
    # interior periodisation is the admissible cheapest one
    tol, n = 1e-8, 150
    th = tc + 0.12 * (2 * rng.uniform(size=n) - 1)
    ph = 1.0 + 0.4 * rng.uniform(size=n)
    pd = O.Plan(th, ph, L, method="direct")
    ps = O.Plan(th, ph, L, tol=tol, method="stretch", nufft="direct")
    assert ps.info["kind_name"] == kind
    assert ps.info["flip"] == flip
    a = O.rand_alm_fast(L, rng)
    assert relmax(ps.synth(a), pd.synth(a)) < 10 * tol
    assert reladj(ps, a, rng.standard_normal(n) + 1j * rng.standard_normal(n)) < 1e-12


def test_stretch_ntrans(O, rng):
    # test_transforms.jl:142-156
    L, B, n = 256, 3, 100
    th = np.deg2rad(20) * np.sqrt(rng.uniform(size=n))
    ph = 2 * np.pi * rng.uniform(size=n)
    p1 = O.Plan(th, ph, L, tol=1e-8, method="stretch", nufft="direct")
    pB = O.Plan(th, ph, L, tol=1e-8, ntrans=B, method="stretch", nufft="direct")
    aB = O.rand_alm_fast(L, rng, ntrans=B)
    fB = pB.synth(aB)
    for b in range(B):
        assert np.abs(fB[:, b] - p1.synth(aB[:, :, b])).max() < 1e-9


def test_dfs_double_is_fold_transpose(O, rng):
    # test_backends.jl:187-200, plus the explicit index formula of src/DFS.jl:19-28
    nth, nph, B = 6, 10, 2
    for s in (0, 1):
        F = rng.standard_normal((nth, nph, B)) + 1j * rng.standard_normal((nth, nph, B))
        Fd = O.dfs_double(F, s)
        G = rng.standard_normal((2 * nth, nph, B)) + 1j * rng.standard_normal((2 * nth, nph, B))
        Gf = O.dfs_fold(G, s)
        assert abs(np.vdot(G, Fd) - np.vdot(Gf, F)) < 1e-12 * abs(np.vdot(G, Fd))
        sg = (-1) ** s
        for i in range(nth):
            assert np.array_equal(Fd[nth + i], sg * np.roll(F[nth - 1 - i], -(nph // 2), axis=0))


def test_fft_matches_numpy(O, rng):
    # radix-4/2/3/5 stages, and the plain-DFT stage for other primes (7, 67, 131)
    for n in (1, 2, 6, 30, 360, 1920, 720, 4, 8, 12, 25, 375, 7, 420, 67, 262):
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        assert relmax(O.fft1(x, -1), np.fft.fft(x)) < 1e-13
        assert relmax(O.fft1(x, +1), np.fft.ifft(x) * n) < 1e-13
    # second-dimension transforms run on blocks of 8 rows: row counts below, at and off a multiple
    for shape in ((12, 20, 2), (8, 45, 1), (3, 64, 3), (45, 96, 1), (17, 50, 2)):
        X = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
        assert relmax(O.fft2(X, -1), np.fft.fft2(X, axes=(0, 1))) < 1e-13
        assert relmax(O.fft2(X, +1, nthreads=3), np.fft.ifft2(X, axes=(0, 1)) * shape[0] * shape[1]) < 1e-13


@pytest.mark.parametrize("tol", [1e-6, 1e-10])
@pytest.mark.parametrize("typ", [1, 2])
def test_es_nufft_vs_direct_sum(O, rng, tol, typ):
    # test_backends.jl:1-19: a fast NUFFT against the reference's direct sum, at 200*tol
    n1, n2, M = 24, 40, 300
    x = rng.uniform(0, np.pi, M)
    y = rng.uniform(0, 2 * np.pi, M)
    if typ == 2:
        c = rng.standard_normal((n1, n2)) + 1j * rng.standard_normal((n1, n2))
        ref = O.nufft_direct(2, n1, n2, +1, x, y, c)
        got = O.nufft_es(2, n1, n2, +1, x, y, c, tol)
    else:
        f = rng.standard_normal(M) + 1j * rng.standard_normal(M)
        ref = O.nufft_direct(1, n1, n2, -1, x, y, f)
        got = O.nufft_es(1, n1, n2, -1, x, y, f, tol)
    assert relmax(got, ref) < 200 * tol


def test_es_fine_size_rule(O):
    # at least 2n and 2w; the next even 5-smooth length up to 4096, then the next of the lengths
    # rocFFT runs in one kernel; past 10240 the next even 5-smooth length again
    cases = {(720, 10): 1440, (1920, 10): 3840, (2048, 10): 4096, (2049, 10): 6144, (2430, 7): 6144,
             (3072, 7): 6144, (3073, 7): 8192, (4097, 10): 9216, (4700, 10): 10000, (5060, 7): 10240,
             (5200, 7): O.next_smooth_even(10400), (3, 10): 20, (60, 10): 120}
    for (n, w), N in cases.items():
        assert O.es_fine_size(n, w) == N, (n, w)
    for n in range(1, 6000, 37):
        N = O.es_fine_size(n, 10)
        assert N >= 2 * n and N >= 20 and N % 2 == 0


@pytest.mark.parametrize("typ", [1, 2])
def test_es_nufft_vs_direct_sum_long_axis(O, rng, typ):
    # an axis whose fine length (6144 for n = 2430) is not the next smooth length: upsampling above
    # 2 keeps the published accuracy
    n1, n2, M, tol = 6, 2430, 120, 1e-8
    assert O.es_fine_size(n2, 10) == 6144
    x = rng.uniform(0, np.pi, M)
    y = rng.uniform(0, 2 * np.pi, M)
    if typ == 2:
        c = rng.standard_normal((n1, n2)) + 1j * rng.standard_normal((n1, n2))
        ref = O.nufft_direct(2, n1, n2, +1, x, y, c)
        got = O.nufft_es(2, n1, n2, +1, x, y, c, tol)
    else:
        f = rng.standard_normal(M) + 1j * rng.standard_normal(M)
        ref = O.nufft_direct(1, n1, n2, -1, x, y, f)
        got = O.nufft_es(1, n1, n2, -1, x, y, f, tol)
    assert relmax(got, ref) < 200 * tol


def test_es_nufft_adjoint_pair(O, rng):
    # type 1 (iflag -1) is the exact adjoint of type 2 (iflag +1)
    n1, n2, M = 30, 44, 500
    x = rng.uniform(0, np.pi, M)
    y = rng.uniform(0, 2 * np.pi, M)
    c = rng.standard_normal((n1, n2)) + 1j * rng.standard_normal((n1, n2))
    f = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    t2 = O.nufft_es(2, n1, n2, +1, x, y, c, 1e-8)[:, 0]
    t1 = O.nufft_es(1, n1, n2, -1, x, y, f, 1e-8)[:, :, 0]
    assert abs(np.vdot(f, t2) - np.vdot(t1, c)) < 1e-13 * abs(np.vdot(f, t2))


def test_planted_mistake_is_caught(O, rng):
    # A dropped phase term (the half-cell shift) in a restated pipeline must fail the Direct gate:
    # shows the gate has teeth. Done on the oracle's stage outputs, in numpy.
    L, tol, n = 128, 1e-8, 60
    th, ph = cap_points(rng, 20, n)
    ps = O.Plan(th, ph, L, tol=tol, method="stretch", nufft="direct")
    pd = O.Plan(th, ph, L, method="direct")
    a = O.rand_alm_fast(L, rng)
    f_ok = ps.synth(a)
    Fhat = ps.last_Fhat()[:, :, 0]
    nth = ps.info["nth"]
    k = np.where(np.arange(nth) < (nth + 1) // 2, np.arange(nth), np.arange(nth) - nth)
    wrong = Fhat * np.exp(1j * k * np.pi / nth)[:, None]  # undo the shift
    f_bad = O.nufft_direct(2, nth, ps.info["nph"], +1, ps_T(O, ps, th), np.mod(ph, 2 * np.pi), wrong)[:, 0]
    fd = pd.synth(a)
    assert relmax(f_ok, fd) < 10 * tol
    assert relmax(f_bad, fd) > 1e3 * tol


def ps_T(O, ps, th):
    g = O.stretch_geometry(0.0, th.max(), ps.L, ps.tol)
    return np.array([O.theta_to_stretch(g, t) for t in th])
