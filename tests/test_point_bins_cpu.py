"""The point map and bin ranges of plan construction (csrc/point_bins.hpp) on the CPU.

The plan-construction kernel runs map_point once per point and the plan build runs bin_ranges on
the per-key counts; here the same header is compiled for the host (tests/cpu_harness) and checked:
- accepted points land where the oracle's stretched coordinate puts them (x1 = T N1 / 2 pi);
- NaN and infinite coordinates are refused (key 0, not counted), finite ones never are;
- the stable counting sort built from the counts gives every accepted point a position < M, and
  bin_ranges refuses counts that do not cover the points exactly (a refused point left in the
  sorted arrays without a count, the condition that once put sorted positions past M).
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import cap_points

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpu_harness", "point_bins_check.cpp")


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("pbc") / "point_bins_check")
    subprocess.run(["/opt/rocm/bin/hipcc", "-O1", "-std=c++17", "-w", "-o", exe, SRC], check=True)
    return exe


def run(exe, text):
    r = subprocess.run([exe], input=text, capture_output=True, text=True, check=True)
    return r.stdout.split("\n")


def point_map_params(O, th, L, tol):
    g = O.stretch_geometry(th.min(), th.max(), L, tol)
    assert g is not None and g.kind == 0  # a north polar cap
    p = O.Plan(th, np.zeros_like(th), L, tol=tol, method="stretch")
    w, _ = O.es_params(tol)
    N1, N2 = p.info["es_n1"], p.info["es_n2"]
    thmax = g.h * (1 + g.epsapo)
    return p, dict(N1=N1, N2=N2, S=17 - w, w=w, thmax=thmax)


def map_points(exe, q, th, ph):
    thmax = float(q["thmax"]).hex()
    head = f"map {q['N1']} {q['N2']} {q['S']} {q['w']} 0 0 1 {thmax} 0 0 {len(th)}\n"
    body = "".join(f"{float(t).hex()} {float(p).hex()}\n" for t, p in zip(th, ph))
    rows = [r.split() for r in run(exe, head + body) if r]
    ok = np.array([int(r[0]) for r in rows], bool)
    key = np.array([int(r[1]) for r in rows])
    b1 = np.array([int(r[2]) for r in rows])
    x1 = np.array([float(r[3]) for r in rows])
    return ok, key, b1, x1


def test_map_matches_oracle_and_refuses_nonfinite(harness, O):
    rng = np.random.default_rng(11)
    L, tol, M = 128, 1e-8, 3000
    th, ph = cap_points(rng, 20, M)
    p, q = point_map_params(O, th, L, tol)
    T, _ = p.points()  # the oracle's stretched coordinate of every point
    ok, key, b1, x1 = map_points(harness, q, th, ph)
    assert ok.all()
    # x1 is the fine-grid coordinate, shifted by one period where the kernel's first cell would be
    # negative (so that every cell index is b1 + t with b1 >= 0)
    want = np.mod(T * (q["N1"] / (2 * np.pi)), q["N1"])
    assert np.allclose(np.mod(x1, q["N1"]), want, rtol=0, atol=1e-9 * q["N1"])
    assert (x1 >= 0).all() and (x1 < 2 * q["N1"]).all() and (b1 >= 0).all() and (b1 < q["N1"]).all()
    assert np.array_equal(b1, np.ceil(x1 - 0.5 * q["w"]).astype(int))
    # refused coordinates mixed in: NaN / inf in theta or phi
    bad_t = [np.nan, np.inf, -np.inf, 0.1, 0.1]
    bad_p = [0.3, 0.3, 0.3, np.nan, np.inf]
    th2 = np.concatenate([th[:20], bad_t])
    ph2 = np.concatenate([ph[:20], bad_p])
    ok2, key2, _, _ = map_points(harness, q, th2, ph2)
    assert ok2[:20].all() and not ok2[20:].any()
    assert (key2[20:] == 0).all()


def test_sorted_positions_below_m(harness, O):
    rng = np.random.default_rng(12)
    L, tol, M = 128, 1e-8, 4000
    th, ph = cap_points(rng, 20, M)
    _, q = point_map_params(O, th, L, tol)
    ok, key, _, _ = map_points(harness, q, th, ph)
    assert ok.all()
    S4 = 4 * q["S"]
    nkeys = ((q["N1"] + S4 - 1) // S4) * ((q["N2"] + S4 - 1) // S4) * 16
    cnt = np.bincount(key, minlength=nkeys)
    out = run(harness, f"ranges {M} {nkeys} " + " ".join(map(str, cnt)) + "\n")[0].split()
    assert out[0] == "ok"
    starts = np.array(out[1:], np.int64)
    assert starts[-1] == M and (np.diff(starts) >= 0).all()
    # stable counting sort: position of point i = start of its key + rank among earlier equal keys
    order = np.argsort(key, kind="stable")
    pos = np.empty(M, np.int64)
    pos[order] = np.arange(M)
    assert pos.max() < M
    assert (pos >= starts[key]).all() and (pos < starts[key + 1]).all()


def test_ranges_refuse_uncounted_points(harness):
    # one point kept in the sorted arrays under key 0 without a count: the ranges would cover M - 1
    # of M points, so the build must refuse instead of launching with them
    cnt = [3, 0, 5, 1]
    out = run(harness, "ranges 10 4 " + " ".join(map(str, cnt)) + "\n")[0]
    assert out.startswith("refused") and "9 of 10" in out
    out = run(harness, "ranges 9 4 " + " ".join(map(str, cnt)) + "\n")[0]
    assert out.startswith("ok")
    out = run(harness, "ranges 3 2 4 -1\n")[0]
    assert out.startswith("refused")
