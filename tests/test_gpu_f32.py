"""Float32 plans and ComplexF32 applies (pfsht_synth_f32 / pfsht_adjoint_f32) through the C ABI.

The reference's Float32 device test (gpu/runtests.jl:122-139): Float32 nodes on a 20 deg cap at
lmax 256, M = 4 (L theta*)^2, a_lm = randn / (1 + l) in ComplexF32, a stretched plan at tol 1e-5; the
device result is ComplexF32 and within 1e-4 of the host plan. Here the host plan is the oracle's FP64
stretched plan on the same (Float32-rounded) nodes. The adjoint is checked the same way and by the
ComplexF32 dot-product identity of test/test_transforms.jl:184-189 (relative 1e-4). A Float32 plan at
tol >= 1e-6 is a mixed plan (FP64 recurrence, FP32 grids); an FP64 plan takes complex64 arrays too,
also from host memory.
"""
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


def ref_cap(rng, L):
    """The reference test's nodes: theta = theta* sqrt(U), phi = 2 pi U, in Float32."""
    ts = np.deg2rad(20.0)
    M = 4 * int(round((L * ts) ** 2))
    th = (ts * np.sqrt(rng.random(M))).astype(np.float32)
    ph = (2 * np.pi * rng.random(M)).astype(np.float32)
    return th, ph


def ref_alm(O, L, rng, B=1):
    """a_lm = randn(ComplexF64) / (1 + l) on the valid modes, rounded to ComplexF32."""
    a = O.rand_alm_fast(L, rng, ntrans=B)
    rows = np.arange(L + 1)[:, None]
    cols = np.arange(2 * L + 1)[None, :]
    l = rows + (cols + 1) // 2
    scale = 1.0 / (1.0 + l)
    if B > 1:
        scale = scale[:, :, None]
    return np.asfortranarray((a * scale).astype(np.complex64))


def test_float32_on_device(torch, P, O, rng):
    L, tol = 256, 1e-5
    th, ph = ref_cap(rng, L)
    M = len(th)
    alm = ref_alm(O, L, rng)
    pd = P.Plan(torch.from_numpy(th), torch.from_numpy(ph), lmax=L, tol=tol, method="stretch")
    assert pd.float32 and pd.precision == "mixed" and pd.info["precision"] == 1
    ad = torch.from_numpy(alm.ravel(order="F").copy()).cuda()
    fd = pd @ ad
    assert fd.dtype == torch.complex64 and fd.is_cuda and fd.numel() == M
    th64, ph64 = th.astype(np.float64), ph.astype(np.float64)
    ph_ = O.Plan(th64, ph64, L, tol=tol, method="stretch", nthreads=NT)
    fh = ph_.synth(alm.astype(np.complex128))
    assert relmax(fd.cpu().numpy().astype(np.complex128), fh) < 1e-4
    # the adjoint, against the oracle and by the ComplexF32 dot-product identity
    fr = (rng.standard_normal(M) + 1j * rng.standard_normal(M)).astype(np.complex64)
    frd = torch.from_numpy(fr).cuda()
    a32 = pd.H @ frd
    assert a32.dtype == torch.complex64
    ah = ph_.adjoint(fr.astype(np.complex128)).ravel(order="F")
    assert relmax(a32.cpu().numpy().astype(np.complex128), ah) < 1e-4
    lhs = torch.vdot(fd, frd).item()
    rhs = torch.vdot(ad, a32).item()
    assert abs(lhs - rhs) / abs(lhs) < 1e-4


def test_float32_end_to_end_direct_and_batched(torch, P, O, rng):
    # test/test_transforms.jl:158-190: Float32 nodes on a Direct plan (L 128, 120 points) against the
    # FP64 evaluation, gate 1e-3 there (Float32 round-off over ~L^2 terms); then a batched stretched
    # plan (ntrans 2) whose columns equal the single applies of an ntrans 1 plan
    L, n = 128, 120
    th = ((20 * np.pi / 180) * np.sqrt(rng.random(n))).astype(np.float32)
    ph = (2 * np.pi * rng.random(n)).astype(np.float32)
    alm = ref_alm(O, L, rng)
    p32 = P.Plan(th, ph, lmax=L, tol=1e-6, method="direct")
    f32 = p32 @ alm  # host complex64 arrays: the host staging of the ComplexF32 path
    assert f32.dtype == np.complex64
    pd = O.Plan(th.astype(np.float64), ph.astype(np.float64), L, method="direct")
    fd = pd.synth(alm.astype(np.complex128))
    assert relmax(f32.astype(np.complex128), fd) < 1e-5
    fr = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    a32 = p32.H @ fr
    assert a32.dtype == np.complex64
    lhs = np.vdot(f32, fr)
    rhs = np.vdot(alm.ravel(order="F"), a32.ravel(order="F"))
    assert abs(lhs - rhs) / abs(lhs) < 1e-4

    L = 256
    th, ph = ref_cap(rng, L)
    M = len(th)
    pB = P.Plan(th, ph, lmax=L, tol=1e-6, ntrans=2, method="stretch")
    p1 = P.Plan(th, ph, lmax=L, tol=1e-6, method="stretch")
    aB = ref_alm(O, L, rng, B=2)
    LB = (L + 1) * (2 * L + 1)
    ad = torch.from_numpy(aB.ravel(order="F").copy()).cuda()
    fB = pB @ ad
    assert fB.dtype == torch.complex64 and fB.numel() == 2 * M
    for b in range(2):
        f1 = p1 @ ad[b * LB:(b + 1) * LB]
        assert relmax(fB[b * M:(b + 1) * M].cpu().numpy(), f1.cpu().numpy()) < 1e-6


def test_fp64_plan_takes_complex64(torch, P, O, rng):
    # an FP64 plan applied to ComplexF32 arrays: the result is the FP64 apply rounded to Float32
    L, tol = 256, 1e-8
    th, ph = ref_cap(rng, L)
    th64, ph64 = th.astype(np.float64), ph.astype(np.float64)
    M = len(th)
    p = P.Plan(th64, ph64, lmax=L, tol=tol, method="stretch")
    assert not p.float32 and p.precision == "fp64"
    alm = ref_alm(O, L, rng)
    f64 = p @ alm.astype(np.complex128)
    f32 = p @ alm
    assert f32.dtype == np.complex64
    assert np.array_equal(f32, f64.astype(np.complex64))
    ad = torch.from_numpy(alm.ravel(order="F").copy()).cuda()
    fd = torch.empty(M, dtype=torch.complex64, device="cuda")
    p.synth(fd, ad)
    assert np.array_equal(fd.cpu().numpy(), f32)
    with pytest.raises(P.ArgumentError):  # mixed element types are refused
        p.synth(torch.empty(M, dtype=torch.complex128, device="cuda"), ad)
