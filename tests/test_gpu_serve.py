"""The serving runtime under a request stream (the serving configuration: caps of 5-30 deg at random
centres, lmax from {256, 512, 1024}, M = 4 (L theta*)^2, ntrans 1-8, tol 1e-8 or 1e-5, synthesis and
adjoint mixed), through the C ABI.

Each served result must equal the same request applied alone on a private plan of the same geometry,
ntrans, tolerance and precision (FP64 plans to round-off, FP32-grid plans within tol). Requests
arriving while others run are packed into shared launches (stats), captured hipGraphs are replayed
with eager work and synchronisations between the replays, a refused cap goes to the Direct leaf, and
the HBM budget evicts idle plans and refuses a plan that cannot fit.
"""
import time

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


# (cap radius deg, lmax, ntrans, tol): a refused cap (Direct), FP32-grid plans (tol 1e-5), FP64 plans
GEOS = [(6.0, 256, 2, 1e-8), (12.0, 256, 3, 1e-5), (20.0, 512, 1, 1e-8), (15.0, 512, 4, 1e-5),
        (25.0, 256, 8, 1e-8), (10.0, 1024, 2, 1e-8)]


def build_geos(torch, P, O):
    rng = np.random.default_rng(55)
    out = []
    for k, (deg, L, B, tol) in enumerate(GEOS):
        tc, pc = np.arccos(rng.uniform(-1, 1)), rng.uniform(0, 2 * np.pi)
        if k == 0:
            tc = 0.0  # a polar cap below 8.25 deg at lmax 256: every periodisation is refused
        M = int(round(4 * (L * np.deg2rad(deg)) ** 2))
        th, ph = centered_cap_points(rng, deg, M, tc, pc)
        prec = "mixed" if tol >= 1e-6 else "fp64"
        alone = P.Plan(th, ph, lmax=L, tol=tol, ntrans=B, method="auto", precision=prec)
        out.append(dict(key=100 + k, th=th, ph=ph, L=L, M=M, B=B, tol=tol, alone=alone,
                        LB=(L + 1) * (2 * L + 1)))
    return out


def rand_input(torch, O, G, direction, seed):
    r = np.random.default_rng(seed)
    if direction == "synth":
        cols = [O.rand_alm_fast(G["L"], r, cl_power=2.0).ravel(order="F") for _ in range(G["B"])]
        return torch.from_numpy(np.concatenate(cols)).cuda()
    n = G["M"] * G["B"]
    return torch.from_numpy(r.standard_normal(n) + 1j * r.standard_normal(n)).cuda()


def alone_result(torch, G, direction, inp):
    n = G["M"] * G["B"] if direction == "synth" else G["LB"] * G["B"]
    out = torch.empty(n, dtype=torch.complex128, device="cuda")
    if direction == "synth":
        G["alone"].synth(out, inp)
    else:
        G["alone"].adjoint(out, inp)
    torch.cuda.synchronize()
    return out


def close(torch, G, got, want):
    e = (got - want).abs().max().item() / want.abs().max().item()
    return e, (e < G["tol"] if G["tol"] >= 1e-6 else e < 1e-11)


def test_serving_stream(torch, P, O):
    geos = build_geos(torch, P, O)
    srv = P.Server(nstreams=3, depth=2, pack=8, graphs=True, hbm_budget_bytes=100 * 10**9)
    for G in geos:
        info, hit = srv.plan(G["key"], G["th"], G["ph"], G["L"], G["tol"], G["B"], "auto")
        assert not hit
        G["method"] = info["method"]
        assert info["precision"] == (1 if G["tol"] >= 1e-6 else 0)
    assert geos[0]["method"] == 1  # the polar 6 deg cap at lmax 256 is refused by the stretch: Direct
    # graph replays, one key and direction at a time, with eager work and synchronisations between
    for G in geos:
        for direction in ("synth", "adjoint"):
            inp = rand_input(torch, O, G, direction, G["key"] * 7 + (direction == "adjoint"))
            want = alone_result(torch, G, direction, inp)
            for rep in range(3):
                out = torch.full_like(want, float("nan"))
                srv.wait(srv.submit(G["key"], direction, inp, out))
                e, ok = close(torch, G, out, want)
                assert ok, (G["key"], direction, rep, e)
                alone_result(torch, G, direction, inp)  # eager work between the replays
    st = srv.stats()
    assert st["graph_launches"] > 0 and st["graph_failures"] == 0 and st["failed"] == 0
    # a Poisson stream over the keys: requests arrive while others run and share launches
    rng = np.random.default_rng(56)
    reqs = []
    for i in range(96):
        G = geos[int(rng.integers(0, len(geos)))]
        direction = "synth" if rng.random() < 0.5 else "adjoint"
        inp = rand_input(torch, O, G, direction, 1000 + i)
        n = G["M"] * G["B"] if direction == "synth" else G["LB"] * G["B"]
        reqs.append((G, direction, inp, torch.full((n,), float("nan"), dtype=torch.complex128, device="cuda")))
    torch.cuda.synchronize()
    gaps = rng.exponential(2e-4, len(reqs))
    launches0, submitted0 = st["launches"], st["submitted"]
    t0 = time.perf_counter()
    arrive = np.cumsum(gaps)
    tickets = []
    for (G, direction, inp, out), ta in zip(reqs, arrive):
        while time.perf_counter() - t0 < ta:
            pass
        tickets.append(srv.submit(G["key"], direction, inp, out))
    for t in tickets:
        srv.wait(t)
    st = srv.stats()
    print(f"\nstream: {len(reqs)} requests in {st['launches'] - launches0} launches; stats {st}", flush=True)
    assert st["failed"] == 0
    worst = 0.0
    for G, direction, inp, out in reqs:
        e, ok = close(torch, G, out, alone_result(torch, G, direction, inp))
        assert ok, (G["key"], direction, e)
        worst = max(worst, e if G["tol"] < 1e-6 else 0.0)
    print(f"FP64 requests: worst relative difference to their alone apply {worst:.2e}", flush=True)
    srv.close()


def test_serving_packs_queued_requests(torch, P, O):
    # one stream with one launch slot: requests of an ntrans 8 key fill a launch each and run one
    # after another; a burst of ntrans 1 requests submitted meanwhile queues behind them and is
    # packed into shared launches, each request still getting its own result
    geos = build_geos(torch, P, O)
    blk, one = geos[4], geos[2]
    assert blk["B"] == 8 and one["B"] == 1
    srv = P.Server(nstreams=1, depth=1, pack=8, graphs=True, hbm_budget_bytes=100 * 10**9)
    for G in (blk, one):
        srv.plan(G["key"], G["th"], G["ph"], G["L"], G["tol"], G["B"], "auto")
    binp = rand_input(torch, O, blk, "synth", 1)
    bouts = [torch.empty(blk["M"] * 8, dtype=torch.complex128, device="cuda") for _ in range(12)]
    inps = [rand_input(torch, O, one, "adjoint", 10 + i) for i in range(8)]
    outs = [torch.full((one["LB"],), float("nan"), dtype=torch.complex128, device="cuda") for _ in range(8)]
    torch.cuda.synchronize()
    st0 = srv.stats()
    bt = [srv.submit(blk["key"], "synth", binp, o) for o in bouts]
    tk = [srv.submit(one["key"], "adjoint", i, o) for i, o in zip(inps, outs)]
    for t in bt + tk:
        srv.wait(t)
    st = srv.stats()
    launches = st["launches"] - st0["launches"]
    print(f"\n20 requests in {launches} launches", flush=True)
    assert launches < 20 and st["failed"] == 0
    for i, o in zip(inps, outs):
        e, ok = close(torch, one, o, alone_result(torch, one, "adjoint", i))
        assert ok, e
    want = alone_result(torch, blk, "synth", binp)
    for o in bouts:
        assert close(torch, blk, o, want)[1]
    srv.close()


def test_serving_admission(torch, P, O):
    geos = build_geos(torch, P, O)[1:4]
    probe = P.Server(nstreams=1, pack=8, hbm_budget_bytes=10**12)
    sizes = [probe.plan(G["key"], G["th"], G["ph"], G["L"], G["tol"], G["B"], "auto")[0]["device_bytes"]
             for G in geos]
    probe.close()
    # admission works on the host estimate of a plan's bytes, an upper bound within a factor of 2
    est = [P.plan_estimate_bytes(G["th"], G["ph"], G["L"], G["tol"], 8, "auto",
                                 "mixed" if G["tol"] >= 1e-6 else "fp64", packed=True) for G in geos]
    for e, a in zip(est, sizes):
        assert a <= e <= 2 * a, (e, a)
    # registered smallest first under a budget of the largest plan's estimate: admitting the largest
    # must evict the others first (they are idle)
    order = np.argsort(est)
    geos = [geos[i] for i in order]
    budget = int(max(est))
    srv = P.Server(nstreams=2, pack=8, hbm_budget_bytes=budget)
    for G in geos:
        srv.plan(G["key"], G["th"], G["ph"], G["L"], G["tol"], G["B"], "auto")
        assert srv.stats()["resident_bytes"] <= budget
    st = srv.stats()
    assert st["evictions"] >= 1 and st["plans"] < 3
    # an evicted key has no plan until it is registered again
    gone = geos[0]
    inp = rand_input(torch, O, gone, "synth", 9)
    out = torch.empty(gone["M"] * gone["B"], dtype=torch.complex128, device="cuda")
    with pytest.raises(P.ArgumentError):
        srv.submit(gone["key"], "synth", inp, out)
    info, hit = srv.plan(gone["key"], gone["th"], gone["ph"], gone["L"], gone["tol"], gone["B"], "auto")
    assert not hit
    srv.wait(srv.submit(gone["key"], "synth", inp, out))
    e, ok = close(torch, gone, out, alone_result(torch, gone, "synth", inp))
    assert ok, e
    # a plan larger than the whole budget is refused before anything is built
    tiny = P.Server(nstreams=1, hbm_budget_bytes=1 << 20)
    with pytest.raises(P.PFSHTError):
        tiny.plan(1, geos[0]["th"], geos[0]["ph"], geos[0]["L"], geos[0]["tol"], geos[0]["B"], "auto")
    assert tiny.stats()["rejected"] == 1
    # bad requests are refused on the host
    with pytest.raises(P.ArgumentError):
        srv.submit(gone["key"], "sideways", inp, out)
    with pytest.raises(P.DimensionMismatch):
        srv.submit(gone["key"], "synth", inp[:-1], out)
    with pytest.raises(P.ArgumentError):
        srv.submit(gone["key"], "synth", inp.cpu(), out.cpu())
    srv.close()
    tiny.close()
