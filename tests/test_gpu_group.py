"""Plan groups (pfsht_group_*): many single-transform plans applied as one, through the C ABI.

Members are drawn as the small-plan batch is (lmax 256, caps of radius U(5, 20) deg around random
centres, M = 4 (L theta*)^2 area-uniform points, tol 1e-8, the planner's method), plus a cap too
small for the stretch, which the planner gives a Direct leaf, a tol 1e-5 member in mixed
precision, and a member whose grid stage takes rocFFT (outside the grouped grid-stage launches). Every member's result equals its own apply (the group runs the same kernels in the same
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
        gf = "library" if i == 2 else "auto"  # a member whose grid stage runs on the member streams
        plans.append(P.Plan(th, ph, lmax=L, tol=tol, method="auto", precision=prec, grid_fft=gf))
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


def test_group_orders_member_applies_across_streams(torch, P, O, rng):
    """A member applied on its own on another stream waits for the group apply that last used its
    work buffers, and a group apply waits for a member's own apply on another stream: a long kernel
    ahead of the first apply keeps the second one's stream from finishing until it ends."""
    plans = members(P, rng, n=4)[:4]
    g = P.Group(plans)
    alms, fs = [], []
    for p in plans:
        a = O.rand_alm_fast(p.lmax, rng, cl_power=2.0)
        alms.append(torch.from_numpy(np.asarray(a).ravel(order="F").copy()).cuda())
        fs.append(torch.from_numpy(rng.standard_normal(p.npts) + 1j * rng.standard_normal(p.npts)).cuda())
    f_ref = [p @ alms[i] for i, p in enumerate(plans)]
    a_ref = [p.H @ fs[i] for i, p in enumerate(plans)]
    fo = [torch.empty_like(f) for f in fs]
    ao = [torch.empty_like(a) for a in alms]
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    for _ in range(4):  # every slot's replay captured before the timed order checks
        g.synth(fo, alms, s1)
        g.adjoint(ao, fs, s1)
    torch.cuda.synchronize()
    done = torch.cuda.Event()

    # group on s1 behind a long kernel, then member 0's own synthesis on s2
    for x in fo:
        x.fill_(np.nan)
    torch.cuda.synchronize()
    with torch.cuda.stream(s1):
        torch.cuda._sleep(int(2e8))
    g.synth(fo, alms, s1)
    x0 = torch.empty_like(fs[0])
    plans[0].synth(x0, alms[0], s2)
    done.record(s2)
    assert not done.query(), "the member's own apply did not wait for the group apply"
    torch.cuda.synchronize()
    assert rel(x0, f_ref[0]) < 1e-14
    for i in range(len(plans)):
        assert rel(fo[i], f_ref[i]) < 1e-14, i

    # member 2's own adjoint on s2 behind a long kernel, then the group adjoint on s1
    y2 = torch.empty_like(alms[2])
    with torch.cuda.stream(s2):
        torch.cuda._sleep(int(2e8))
    plans[2].adjoint(y2, fs[2], s2)
    g.adjoint(ao, fs, s1)
    done.record(s1)
    assert not done.query(), "the group apply did not wait for the member's own apply"
    torch.cuda.synchronize()
    assert rel(y2, a_ref[2]) < 1e-14
    for i in range(len(plans)):
        assert rel(ao[i], a_ref[i]) < 1e-14, i
