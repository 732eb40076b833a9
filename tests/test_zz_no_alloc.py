"""The allocation-free adjoint ring pass (pfsht_alm_from_rings_noalloc) against pfsht_alm_from_rings,
through the C ABI: the same kernels with their scratch in the library's static device array, so the
results must be bit-identical. Two calls in a row, the second with a smaller engine (its scratch is a
smaller part of the same array); a spin engine; an engine whose scratch exceeds the array is refused."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as T
    assert T.cuda.is_available(), "a GPU is required for the gpu-marked tests"
    return T


def cap_rings(n, deg):
    return np.deg2rad(deg) * (np.arange(n) + 0.5) / n


def one(torch, P, L, n, B, spin=0, seed=0):
    R = P.Rings(cap_rings(n, 20), L, 1e-10, ntrans=B, spin=spin)
    r = np.random.default_rng(seed)
    G = torch.from_numpy(r.standard_normal(n * (2 * L + 1) * B) + 1j * r.standard_normal(n * (2 * L + 1) * B)).cuda()
    LB = (L + 1) * (2 * L + 1) * B
    want = torch.full((LB,), float("nan"), dtype=torch.complex128, device="cuda")
    got = torch.full((LB,), float("nan"), dtype=torch.complex128, device="cuda")
    R.alm_from_rings(want, G)
    R.alm_from_rings_noalloc(got, G)
    torch.cuda.synchronize()
    assert torch.equal(got, want), (L, n, B, spin)
    return R


def test_noalloc_matches_alm_from_rings(torch):
    import pfsht as P
    one(torch, P, 512, 160, 2, seed=1)   # MFMA adjoint: ring buffer only
    one(torch, P, 256, 80, 1, seed=2)    # smaller, single transform: ring buffer and partials
    one(torch, P, 128, 50, 3, spin=2, seed=3)
    one(torch, P, 512, 160, 1, seed=4)   # larger again after the smaller call
    big = P.Rings(cap_rings(360, 20), 2048, 1e-8)
    G = torch.zeros(360 * 4097, dtype=torch.complex128, device="cuda")
    a = torch.zeros(2049 * 4097, dtype=torch.complex128, device="cuda")
    with pytest.raises(P.ArgumentError):
        big.alm_from_rings_noalloc(a, G)
