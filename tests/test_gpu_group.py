"""Plan groups (pfsht_group_*): many single-transform plans applied as one, through the C ABI.

Members are drawn as the small-plan batch is (lmax 256, caps of radius U(5, 20) deg around random
centres, M = 4 (L theta*)^2 area-uniform points, tol 1e-8, the planner's method), plus a cap too
small for the stretch, which the planner gives a Direct leaf, and a tol 1e-5 member in mixed
precision. Every member's result equals its own apply (the group runs the same kernels in the same
order: the ring passes grouped into one launch, the rest on concurrent streams) in both directions,
on repeated calls with new inputs (the hipGraph replays), with eager work of the members' own in
between; the adjoint identity holds per member.
"""
import numpy as np
import pytest

from conftest import centered_cap_points

pytestmark = pytest.mark.gpu


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


def members(P, rng, n=8, L=256):
    plans = []
    for i in range(n):
        deg = rng.uniform(5.0, 20.0)
        tc, pc = np.arccos(rng.uniform(-1.0, 1.0)), rng.uniform(0.0, 2 * np.pi)
        M = int(round(4 * (L * np.deg2rad(deg)) ** 2))
        th, ph = centered_cap_points(rng, deg, M, tc, pc)
        tol, prec = (1e-5, "mixed") if i == 1 else (1e-8, "fp64")
        plans.append(P.Plan(th, ph, lmax=L, tol=tol, method="auto", precision=prec))
    th, ph = centered_cap_points(rng, 0.5, 60, 1.0, 2.0)  # below the stretch's reach: Direct
    plans.append(P.Plan(th, ph, lmax=L, tol=1e-8, method="auto"))
    return plans


def rel(a, b):
    return (a - b).abs().max().item() / max(b.abs().max().item(), 1e-300)


def test_group_matches_members(torch, P, O, rng):
    plans = members(P, rng)
    assert plans[-1].algorithm == "direct" and plans[1].precision == "mixed"
    assert sum(p.algorithm == "stretch" for p in plans) >= 6
    g = P.Group(plans)
    for call in range(4):
        alms, fs = [], []
        for p in plans:
            a = O.rand_alm_fast(p.lmax, rng, cl_power=2.0)
            alms.append(torch.from_numpy(np.asarray(a).ravel(order="F").copy()).cuda())
            fs.append(torch.from_numpy(rng.standard_normal(p.npts) + 1j * rng.standard_normal(p.npts)).cuda())
        fo = [torch.full((p.npts,), np.nan, dtype=torch.complex128, device="cuda") for p in plans]
        ao = [torch.full(((p.lmax + 1) * (2 * p.lmax + 1),), np.nan, dtype=torch.complex128, device="cuda")
              for p in plans]
        g.synth(fo, alms)
        g.adjoint(ao, fs)
        torch.cuda.synchronize()
        for i, p in enumerate(plans):
            f1 = p @ alms[i]
            a1 = p.H @ fs[i]
            es, ea = rel(fo[i], f1), rel(ao[i], a1)
            assert es < 1e-14 and ea < 1e-14, (call, i, p.algorithm, es, ea)
            # the adjoint identity <Y a, f> = <a, Y' f>: to round-off in FP64, to the plan's tolerance
            # with FP32 grids
            lhs = torch.vdot(fs[i], fo[i]).item()
            rhs = torch.vdot(ao[i], alms[i]).item()
            bound = 1e-10 if p.precision == "fp64" else p.tol
            assert abs(lhs - rhs) <= bound * abs(lhs), (call, i, abs(lhs - rhs) / abs(lhs))
        print(f"call {call}: {len(plans)} members equal to their own applies", flush=True)


def test_group_arguments(torch, P, rng):
    th, ph = centered_cap_points(rng, 10.0, 5000, 0.5, 0.5)
    p2 = P.Plan(th, ph, lmax=256, tol=1e-8, ntrans=2, method="stretch")
    p1 = P.Plan(th, ph, lmax=256, tol=1e-8, method="stretch")
    with pytest.raises(P.ArgumentError):  # ntrans > 1 members are refused (PFSHT_E_UNSUPPORTED)
        P.Group([p1, p2])
    with pytest.raises(P.ArgumentError):
        P.Group([p1, p1])
    g = P.Group([p1])
    a = torch.zeros(257 * 513, dtype=torch.complex128, device="cuda")
    with pytest.raises(P.DimensionMismatch):
        g.synth([torch.zeros(4999, dtype=torch.complex128, device="cuda")], [a])
    with pytest.raises(P.DimensionMismatch):
        g.synth([], [])
