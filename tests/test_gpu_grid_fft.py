"""The grid stage's LDS kernels (grid_fft_kernels.hip: the colatitude, fine-column and fine-row FFTs
inside workgroups, fused with the window, periodisation, deconvolution and transposes) against the
rocFFT path of the same plan geometry (grid_fft="library"), through the C ABI, on every periodisation
kind, spin, batched columns, mixed precision, and fine lengths that take each workgroup size of the
LDS transforms (128 ... 1024 threads; rows past 8192 values split in two, here 10000). The two paths
compute the same transform with different FFT algorithms, so they agree to round-off: 1e-12 relative
in FP64, 2e-5 in mixed precision (FP32 grids). The widest case is also held to the oracle's Direct
sum at 10 tol."""
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


def coeffs(O, L, rng, spin, B):
    cols = []
    for _ in range(B):
        if spin:
            a = np.zeros((L + 1, 2 * L + 1), complex, order="F")
            m = O.spin_valid_mask(L, spin)
            a[m] = rng.standard_normal(m.sum()) + 1j * rng.standard_normal(m.sum())
        else:
            a = O.rand_alm_fast(L, rng, cl_power=2.0)
        cols.append(np.asarray(a).ravel(order="F"))
    return np.concatenate(cols)


def both(P, th, ph, **kw):
    lds = P.Plan(th, ph, grid_fft="lds", **kw)
    lib = P.Plan(th, ph, grid_fft="library", **kw)
    assert lds.info["grid_fft"] == 1 and lib.info["grid_fft"] == 2
    for k in ("nth", "nph", "mmax", "es_n1", "es_n2", "es_rows", "kind"):
        assert lds.info[k] == lib.info[k], k
    return lds, lib


def compare(torch, O, lds, lib, L, M, B, spin, rng, bound):
    a = torch.from_numpy(coeffs(O, L, rng, spin, B)).cuda()
    f1 = torch.empty(M * B, dtype=torch.complex128, device="cuda")
    f2 = torch.empty_like(f1)
    lds.synth(f1, a)
    lib.synth(f2, a)
    assert (f1 - f2).abs().max().item() <= bound * f2.abs().max().item()
    g = torch.from_numpy(rng.standard_normal(M * B) + 1j * rng.standard_normal(M * B)).cuda()
    a1 = torch.empty(a.numel(), dtype=torch.complex128, device="cuda")
    a2 = torch.empty_like(a1)
    lds.adjoint(a1, g)
    lib.adjoint(a2, g)
    assert (a1 - a2).abs().max().item() <= bound * a2.abs().max().item()
    torch.cuda.synchronize()


CASES = {
    # name: (points, lmax, tol, spin, ntrans, keywords)
    "pole": ("cap20", 256, 1e-8, 0, 1, {}),
    "interior": ("band", 256, 1e-8, 0, 1, {}),
    "south_flip": ("south20", 256, 1e-10, 0, 1, {}),
    "spin2": ("cap20", 128, 1e-8, 2, 1, {}),
    "batched3": ("cap20", 200, 1e-8, 0, 3, {}),
    "mixed": ("cap20", 256, 1e-5, 0, 2, {"precision": "mixed"}),
    "wide_nt512": ("cap20", 256, 1e-8, 0, 1, {"oversample_phi": 12.0}),
}


def points(name, rng, M):
    if name == "cap20":
        return cap_points(rng, 20, M)
    if name == "south20":
        return centered_cap_points(rng, 20, M, np.pi, 0.5)
    return centered_cap_points(rng, 12, M, np.pi / 2, 1.0)


@pytest.mark.parametrize("case", sorted(CASES))
def test_lds_grid_matches_library(torch, P, O, rng, case):
    where, L, tol, spin, B, kw = CASES[case]
    M = 4000
    th, ph = points(where, rng, M)
    lds, lib = both(P, th, ph, lmax=L, tol=tol, spin=spin, ntrans=B, method="stretch", **kw)
    if case == "interior":
        assert lds.info["kind_name"] == "interior"
    if case == "wide_nt512":
        assert lds.info["es_n2"] > 4096  # a 512-thread row transform
    bound = 2e-5 if kw.get("precision") == "mixed" else 1e-12
    compare(torch, O, lds, lib, L, M, B, spin, rng, bound)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_lds_grid_longest_rows(torch, P, O, rng, precision):
    # lmax 1024 with oversample_phi 4.6: fine rows of 10000 values (156 KiB of FP64 LDS, 1024 threads)
    L, M = 1024, 3000
    tol = 1e-8 if precision == "fp64" else 1e-5
    th, ph = cap_points(rng, 20, M)
    lds, lib = both(P, th, ph, lmax=L, tol=tol, method="stretch", precision=precision, oversample_phi=4.6)
    assert 8192 < lds.info["es_n2"] <= 10240
    compare(torch, O, lds, lib, L, M, 1, 0, rng, 1e-12 if precision == "fp64" else 2e-5)
    if precision == "fp64":
        a = O.rand_alm_fast(L, rng, cl_power=2.0)
        sub = rng.choice(M, 200, replace=False)
        pd = O.Plan(th[sub], ph[sub], L, method="direct", nthreads=NT)
        assert relmax((lds @ a)[sub], pd.synth(a)) < 10 * tol
        # past a workgroup's LDS (rows of 14400 FP64 values): "lds" is refused, "auto" takes rocFFT
        with pytest.raises(P.ArgumentError):
            P.Plan(th, ph, lmax=L, tol=tol, method="stretch", oversample_phi=7.0, grid_fft="lds")
        wide = P.Plan(th, ph, lmax=L, tol=tol, method="stretch", oversample_phi=7.0)
        assert wide.info["es_n2"] > 10240 and wide.info["grid_fft"] == 2


def test_lds_grid_partial_columns(torch, P, O, rng):
    # a packed plan applying fewer columns than its ntrans: the LDS kernels run nb transforms
    L, M, B = 200, 3000, 4
    th, ph = cap_points(rng, 20, M)
    lds, lib = both(P, th, ph, lmax=L, tol=1e-8, ntrans=B, method="stretch", packed=True)
    LB = (L + 1) * (2 * L + 1)
    for nb in (1, 3):
        alms = [torch.from_numpy(coeffs(O, L, rng, 0, 1)).cuda() for _ in range(nb)]
        f1 = [torch.empty(M, dtype=torch.complex128, device="cuda") for _ in range(nb)]
        f2 = [torch.empty(M, dtype=torch.complex128, device="cuda") for _ in range(nb)]
        lds.synth_cols(f1, alms)
        lib.synth_cols(f2, alms)
        for x, y in zip(f1, f2):
            assert (x - y).abs().max().item() <= 1e-12 * y.abs().max().item()
        gs = [torch.from_numpy(rng.standard_normal(M) + 1j * rng.standard_normal(M)).cuda() for _ in range(nb)]
        o1 = [torch.empty(LB, dtype=torch.complex128, device="cuda") for _ in range(nb)]
        o2 = [torch.empty(LB, dtype=torch.complex128, device="cuda") for _ in range(nb)]
        lds.adjoint_cols(o1, gs)
        lib.adjoint_cols(o2, gs)
        for x, y in zip(o1, o2):
            assert (x - y).abs().max().item() <= 1e-12 * y.abs().max().item()
    torch.cuda.synchronize()
