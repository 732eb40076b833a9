"""The plan-time tuner of the ring passes' launch shape, through the C ABI.

A stretched plan times candidate shapes on its own ring engine, keeps the fastest and caches it per
(architecture, lmax, rings, largest order, ntrans): a second plan of the shape times nothing and gets
the same shape. Every shape the tuner can choose (segment targets 2, 3, 4, 6; one or two tiles per
wavefront of the single-transform and of the MFMA adjoint) is forced through the cache and must give
the default shape's results to round-off (only the summation order of the segments differs).
"""
import numpy as np
import pytest

from conftest import cap_points

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as T
    assert T.cuda.is_available(), "a GPU is required for the gpu-marked tests"
    return T


@pytest.fixture
def P():
    import pfsht
    pfsht.lib()
    pfsht.tune_clear()
    yield pfsht
    pfsht.tune_clear()


def rel(a, b):
    return (a - b).abs().max().item() / b.abs().max().item()


def test_tuned_and_cached(torch, P, rng):
    th, ph = cap_points(rng, 20, 20000)
    p1 = P.Plan(th, ph, lmax=512, tol=1e-8, method="stretch")
    rows = p1.tune_table()
    assert len(rows) >= 4 and all(r["ms"] > 0 for r in rows)
    key = p1.tune_key()
    assert key[0].startswith("gfx") and key[1] == 512 and key[2] == p1.info["nth_half"] and key[4] == 1
    got = P.tune_get(*key)
    assert got is not None
    shape = (p1.info["seg_blocks"], p1.info["adj_tiles"], p1.info["mfma_tiles"])
    assert shape == (got["seg_blocks"], got["adj_tiles"], got["mfma_tiles"])
    assert any((r["seg_blocks"], r["adj_tiles"], r["mfma_tiles"]) == shape for r in rows)
    print(f"\nlmax 512: {rows} -> {shape}", flush=True)
    # the same ring set again: the cached shape, nothing timed
    p2 = P.Plan(th, ph, lmax=512, tol=1e-8, method="stretch")
    assert p2.tune_table() == []
    assert (p2.info["seg_blocks"], p2.info["adj_tiles"], p2.info["mfma_tiles"]) == shape
    # tune = -1 on an empty cache: the default shape, nothing timed
    P.tune_clear()
    p3 = P.Plan(th, ph, lmax=512, tol=1e-8, method="stretch", tune=-1)
    assert p3.tune_table() == [] and p3.info["seg_blocks"] == 3 and P.tune_size() == 0


@pytest.mark.parametrize("L,deg,B", [(512, 20, 1), (512, 20, 4), (1024, 30, 2)])
def test_every_shape_matches_default(torch, P, O, rng, L, deg, B):
    M = 8000
    th, ph = cap_points(rng, deg, M)
    base = P.Plan(th, ph, lmax=L, tol=1e-8, ntrans=B, method="stretch", tune=-1)
    key = base.tune_key()
    LB = (L + 1) * (2 * L + 1)
    a = torch.from_numpy(np.asarray(O.rand_alm_fast(L, rng, ntrans=B, cl_power=2.0)).ravel(order="F").copy()).cuda()
    f = torch.from_numpy(rng.standard_normal(M * B) + 1j * rng.standard_normal(M * B)).cuda()
    f0 = torch.empty(M * B, dtype=torch.complex128, device="cuda")
    a0 = torch.empty(LB * B, dtype=torch.complex128, device="cuda")
    base.synth(f0, a)
    base.adjoint(a0, f)
    mfma = base.info["mfma_tiles"] > 0
    shapes = [(per, 0, 0) for per in (2, 3, 4, 6)]
    shapes += [(3, 1, 0)] if not mfma else [(3, 0, 1), (3, 0, 2)]
    seen = set()
    for per, atg, mtg in shapes:
        P.tune_clear()
        P.tune_put(*key, seg_blocks=per, adj_tiles=atg, mfma_tiles=mtg)
        p = P.Plan(th, ph, lmax=L, tol=1e-8, ntrans=B, method="stretch", tune=-1)
        got = (p.info["seg_blocks"], p.info["adj_tiles"], p.info["mfma_tiles"])
        seen.add(got)
        assert got[0] == per
        fs = torch.empty_like(f0)
        aa = torch.empty_like(a0)
        p.synth(fs, a)
        p.adjoint(aa, f)
        torch.cuda.synchronize()
        print(f"L {L} B {B} shape {got}: synth {rel(fs, f0):.2e} adjoint {rel(aa, a0):.2e}", flush=True)
        assert rel(fs, f0) < 1e-12 and rel(aa, a0) < 1e-12, got
        del p
    assert len(seen) >= 5
