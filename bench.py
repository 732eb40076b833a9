#!/usr/bin/env python3
"""Plain-apply throughput of the stretched-leaf path on one MI355X.

Workload: 20 deg polar cap, lmax = 2048, M = 2.0e6 area-uniform points, tol = 1e-8, ntrans = 1,
FP64 throughout; alm complex with Re/Im ~ N(0, C_l/2), C_l = (l+1)^-2 (a_l0 real); f complex
N(0, 1). One step = one synthesis mul!(f, plan, alm) plus one adjoint mul!(alm, plan', f), inputs
resident in HBM. `value` counts applies (either direction) per second. The same line carries a short
served leg (8 recurring geometries of the serving workload through the serving runtime: p50/p99
latency, served against one-at-a-time applies) and the CPU baseline (serial and threaded).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-serve]

Prints one JSON line (the last line of stdout). Progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "partialfsht.jl_amd"))

PEAK_FP64_TFLOPS = 78.6  # MI355X dense FP64 (vector and matrix)
PEAK_HBM_GBS = 8000.0


def log(msg):
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def cap_points(rng, deg, M):
    ts = np.deg2rad(deg)
    th = np.arccos(rng.uniform(np.cos(ts), 1.0, M))
    ph = rng.uniform(0.0, 2 * np.pi, M)
    return th, ph


def rand_alm(L, rng, cl_power=2.0):
    rows = np.arange(L + 1)[:, None]
    cols = np.arange(2 * L + 1)[None, :]
    l = rows + (cols + 1) // 2
    z = rng.standard_normal((L + 1, 2 * L + 1)) + 1j * rng.standard_normal((L + 1, 2 * L + 1))
    z *= np.sqrt(0.5 * (l + 1.0) ** (-cl_power))
    z[:, 0] = z[:, 0].real
    z[l > L] = 0
    return np.asfortranarray(z)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(th, ph, L, tol, alm, f, nthreads):
    """The oracle's restatement of the reference host path (test infrastructure, timed as the
    baseline only) on this host's cores, by the reference's own protocol (benchmark/calibrate.jl:
    56-63, 81-101): a serial leg and a threaded leg, each synthesis and adjoint timed best of 3 after
    one warm-up, at the full workload. The NUFFT is the oracle's ES-kernel stand-in for FINUFFT (not
    vendored); both FFT stages use the oracle's Stockham mixed-radix FFT, threaded in the threaded
    leg. Stage times (ring / grid incl. the pole FFT / NUFFT incl. the fine FFT) are those of the
    best run."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    O.build()

    def best(p, fn, x):
        fn(x)  # warm-up
        t, st = float("inf"), None
        for _ in range(3):
            t0 = time.perf_counter()
            fn(x)
            dt = time.perf_counter() - t0
            if dt < t:
                t, st = dt, p.last_times()
        return t, st

    legs = {}
    for name, nt in (("serial", 1), ("threaded", nthreads)):
        t0 = time.perf_counter()
        p = O.Plan(th, ph, L, tol=tol, method="stretch", nufft="es", nthreads=nt)
        t_plan = time.perf_counter() - t0
        ts, ss = best(p, p.synth, alm)
        ta, sa = best(p, p.adjoint, f)
        W = p.info["ring_work"]
        legs[name] = {"threads": nt, "synth_s": ts, "adjoint_s": ta, "applies_s": 2.0 / (ts + ta),
                      "plan_s": t_plan,
                      "synth_stage_s": {"ring": ss[0], "grid": ss[1], "nufft": ss[2]},
                      "adjoint_stage_s": {"ring": sa[0], "grid": sa[1], "nufft": sa[2]},
                      "ns_per_triple_ring_synth": ss[0] * 1e9 / W}
        log(f"cpu {name} ({nt} threads): synth {ts * 1e3:.1f} ms, adjoint {ta * 1e3:.1f} ms "
            f"(ring {ss[0] * 1e3:.1f} / grid {ss[1] * 1e3:.1f} / nufft {ss[2] * 1e3:.1f} ms)")
        del p
    thr = legs["threaded"]
    return {"value": thr["applies_s"], "unit": "applies/s", "cores": nthreads, "kind": "port",
            "sample": f"the full workload (lmax {L}, M {len(th)}), one synthesis + one adjoint, best of 3 "
                      f"after a warm-up, at 1 thread and at {nthreads} threads; value = the threaded leg",
            "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
            "cpus_allowed": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None,
            "serial": legs["serial"], "threaded": thr,
            "published_reference": {"serial_synth_s": 0.538, "threaded_synth_s": 0.245,
                                    "ring_ns_per_triple": 1.45,
                                    "source": "README.md:44,55: lmax 2048, M 2e6, tol 1e-8, FINUFFT, 4 "
                                              "physical cores at -t4"}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", choices=["plain", "batched", "cap30", "equatorial", "cap15at60", "spin", "twocaps",
                                         "serve", "tune", "smallplans"],
                    default="plain",
                    help="plain: the 20 deg cap at lmax 2048, M 2e6 (default, the headline workload); "
                         "batched: the same cap at lmax 1024, M 5.1e5, ntrans 16, against one-at-a-time "
                         "applies of the same plan geometry; cap30: 30 deg polar cap, lmax 2048, "
                         "M = 4 (L theta*)^2, tol 1e-5; equatorial / cap15at60: interior caps (30 deg "
                         "around the equator, 15 deg around colatitude 60 deg) at lmax 2048, tol 1e-8; "
                         "these three also report the error against the Direct leaf on 1e4 points; "
                         "serve: a Poisson request stream over 32 recurring geometries through the "
                         "serving runtime (plan cache, HBM admission, 4 streams), p50/p99 latency; "
                         "tune: the ring-pass shapes the plan-time tuner picks at lmax 256 / 1024 / 2048; "
                         "smallplans: 64 small plans (lmax 256) as one plan group against one at a time")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-serve", action="store_true", help="plain: skip the short served leg")
    ap.add_argument("--lmax", type=int, default=2048)
    ap.add_argument("--npts", type=int, default=2_000_000)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64",
                    help="cap30 / equatorial / cap15at60 only: mixed = FP64 ring recurrence with the grid "
                         "buffers and FFTs in FP32 (needs tol >= 1e-6)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 256 MiB device-to-device copy three times (HBM counter calibration "
                         "for rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        x = torch.empty(1 << 28, dtype=torch.uint8, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": 1 << 28, "launches": 3}))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path (collectives are out of "
              "scope); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.config == "batched":
        return run_batched(args)
    if args.config in ("cap30", "equatorial", "cap15at60", "spin", "twocaps"):
        return run_geometry(args)
    if args.config == "serve":
        return run_serve(args)
    if args.config == "tune":
        return run_tune(args)
    if args.config == "smallplans":
        return run_smallplans(args)
    return run_plain(args)


def run_plain(args):
    import torch
    import pfsht
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    L, tol, M = args.lmax, 1e-8, args.npts
    rng = np.random.default_rng(2048)
    th, ph = cap_points(rng, 20, M)
    alm = rand_alm(L, rng)
    f = rng.standard_normal(M) + 1j * rng.standard_normal(M)

    log(f"plan lmax={L} M={M} tol={tol}")
    t0 = time.perf_counter()
    plan = pfsht.Plan(th, ph, lmax=L, tol=tol, method="stretch")
    torch.cuda.synchronize()
    t_plan = time.perf_counter() - t0
    info = plan.info
    log(f"plan built in {t_plan:.2f}s: nth={info['nth']} nph={info['nph']} mmax={info['mmax']} "
        f"ring_work={info['ring_work']} fine={info['es_n1']}x{info['es_n2']} w={info['es_w']}")
    # the first plan of a process pays HIP start-up and rocFFT's run-time kernel compilation; a second
    # plan on new points of the same sizes shows the plan construction itself
    th2, ph2 = cap_points(np.random.default_rng(4096), 20, M)
    t0 = time.perf_counter()
    plan2 = pfsht.Plan(th2, ph2, lmax=L, tol=tol, method="stretch")
    torch.cuda.synchronize()
    t_plan_warm = time.perf_counter() - t0
    del plan2
    log(f"second plan (new points) built in {t_plan_warm:.3f}s")

    alm_d = pfsht.to_device(alm)
    f_d = pfsht.to_device(f)
    f_out = torch.empty(M, dtype=torch.complex128, device="cuda")
    a_out = torch.empty(alm.size, dtype=torch.complex128, device="cuda")
    stream = torch.cuda.current_stream()

    for _ in range(args.warmup):
        plan.synth(f_out, alm_d, stream)
        plan.adjoint(a_out, f_d, stream)
    torch.cuda.synchronize()
    log(f"warmup done ({args.warmup} steps)")

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plan.synth(f_out, alm_d, stream)
        plan.adjoint(a_out, f_d, stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    # per-stage device times from HIP events on the apply stream, in a separate pass (recording
    # events costs device time, so the timed steps above run without them); also the per-apply
    # latency distribution from events around each apply
    plan.set_timing(True)
    plan.stage_times("synth", reset=True)
    plan.stage_times("adjoint", reset=True)
    ev = []
    for _ in range(args.steps):
        e = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
        e[0].record(stream)
        plan.synth(f_out, alm_d, stream)
        e[1].record(stream)
        plan.adjoint(a_out, f_d, stream)
        e[2].record(stream)
        ev.append(e)
    torch.cuda.synchronize()
    plan.set_timing(False)
    lat_s = np.array([e[0].elapsed_time(e[1]) for e in ev])
    lat_a = np.array([e[1].elapsed_time(e[2]) for e in ev])
    ms_step = elapsed / args.steps * 1e3
    applies_s = 2 * args.steps / elapsed
    ss = plan.stage_times("synth")
    sa = plan.stage_times("adjoint")
    log(f"timed {args.steps} steps: {ms_step:.3f} ms/step, {applies_s:.1f} applies/s")

    W = info["ring_work"]
    nu_b = 2 * 16 * M + 16 * info["es_n1"] * info["es_n2"] + 4 * M  # points + f, fine grid, perm
    # the grid stage (LDS column + row kernels): the ring columns, the column/row hand-off written and
    # read once, the stored fine rows
    ncol, nrow = 2 * info["mmax"] + 1, info["es_rows"]
    grid_b = 16 * (ncol * info["nth_half"] + 2 * ncol * nrow + nrow * info["es_n2"])

    def both(d, *keys):
        return (sum(d[k][0] for k in keys), max(d[k][1] for k in keys))

    stages = {}
    for name, d, flops, bytes_ in (
            ("ring_synth", ss["ring"], 12.0 * W, None),
            ("ring_adjoint", sa["ring"], 12.0 * W, None),
            ("nufft_interp", ss["nufft"], None, nu_b),
            ("nufft_spread", sa["nufft"], None, nu_b),
            ("grid_synth", both(ss, "grid_fft", "fine_fft"), None, grid_b),
            ("grid_adjoint", both(sa, "grid_fft", "fine_fft"), None, grid_b)):
        ms_avg = d[0] / max(d[1], 1)
        e = {"ms": ms_avg}
        if flops:
            e["tflops"] = flops / (ms_avg * 1e-3) / 1e12
            e["frac_fp64_peak"] = e["tflops"] / PEAK_FP64_TFLOPS
        if bytes_:
            e["gbs"] = bytes_ / (ms_avg * 1e-3) / 1e9
            e["frac_hbm_peak"] = e["gbs"] / PEAK_HBM_GBS
        stages[name] = e
    per_stage = {"synth": {k: v[0] / max(v[1], 1) for k, v in ss.items()},
                 "adjoint": {k: v[0] / max(v[1], 1) for k, v in sa.items()}}
    # the dominant kernel of the step: the adjoint ring pass (ring_adj_kernel), timed alone
    ring = stages["ring_adjoint"]
    traffic, traffic_src = None, None
    tf = os.path.join(ROOT, "profiles", "pmc_traffic_lmax2048_M2e6.json")
    if os.path.exists(tf) and L == 2048 and M == 2_000_000:
        t = json.load(open(tf))
        k = t.get("kernels", {}).get("ring_adj_kernel")
        if k:
            traffic, traffic_src = k["read_bytes"] + k["write_bytes"], "profiles/pmc_traffic_lmax2048_M2e6.json"
    roofline = {"kernel": "ring_adj_kernel (adjoint ring pass)", "bound": "valu",
                "achieved": ring["tflops"], "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                "frac": ring["tflops"] / PEAK_FP64_TFLOPS, "traffic": traffic,
                "traffic_source": traffic_src,
                "algorithmic": f"12 FP64 flops per retained (l,m,ring) triple x {W} triples per launch",
                "bound_note": "FP64 vector FMA issue (v_fma_f64 on the VALU, one transform: no matrix "
                              "operand to fill); MI355X's FP64 vector and FP64 matrix peaks are both "
                              "78.6 TFLOP/s"}

    served = None
    if not args.no_serve:
        log("served leg: 8 geometries, 160 requests")
        served = serve_leg(8, 160)

    cpu = None
    if not args.no_cpu:
        nt = min(16, os.cpu_count() or 1)
        log(f"cpu baseline (oracle, serial and {nt} threads)")
        cpu = cpu_baseline(th, ph, L, tol, alm, f, nt)
        cpu["gpu_over_threaded"] = applies_s / cpu["value"]
        cpu["gpu_over_serial"] = applies_s / cpu["serial"]["applies_s"]
        log(f"cpu baseline: {cpu['value']:.3f} applies/s threaded, {cpu['serial']['applies_s']:.3f} serial")

    out = {"metric": "applies/s at lmax=2048 tol=1e-8 (synth+adjoint); served p50/p99 latency",
           "value": applies_s, "unit": "applies/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"plain apply, 20deg polar cap, lmax={L}, M={M} area-uniform, "
                                  f"tol={tol}, ntrans=1, one synthesis + one adjoint per step",
                      "lmax": L, "npts": M, "tol": tol, "ntrans": 1,
                      "nth": info["nth"], "nph": info["nph"], "mmax": info["mmax"],
                      "ring_work": W, "fine_grid": [info["es_n1"], info["es_n2"]], "fine_rows_stored": info["es_rows"],
                      "es_w": info["es_w"]},
           "triples_per_s": W * 2 * args.steps / elapsed,
           "apply_latency_ms": {"synthesis": {"p50": float(np.percentile(lat_s, 50)),
                                              "p99": float(np.percentile(lat_s, 99))},
                                "adjoint": {"p50": float(np.percentile(lat_a, 50)),
                                            "p99": float(np.percentile(lat_a, 99))},
                                "note": "device time of single applies on an idle GPU, stage events on"},
           "served": served,
           "plan_s": t_plan,
           "plan_s_warm": t_plan_warm,
           "stage_ms": per_stage,
           "stages": stages,
           "roofline": roofline,
           "cpu_baseline": cpu}
    print(json.dumps(out))


def centered_cap(rng, deg, M, theta_c, phi_c=0.0):
    """Area-uniform points on the cap of radius deg degrees around (theta_c, phi_c)."""
    th, ph = cap_points(rng, deg, M)
    x, y, z = np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)
    c, s = np.cos(theta_c), np.sin(theta_c)
    x, z = c * x + s * z, -s * x + c * z
    return np.arccos(np.clip(z, -1.0, 1.0)), np.mod(np.arctan2(y, x) + phi_c, 2 * np.pi)


def run_geometry(args):
    """Larger caps, one plan each, synthesis + adjoint per step, and the delivered error against the
    device Direct leaf on a 1e4-point random subset: synthesis on the subset, and the adjoint of a
    field supported on the subset (for which Direct is exact and cheap)."""
    import torch
    import pfsht
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    L = 2048
    geo = {"cap30": (30.0, 0.0, 1e-5, "30deg polar cap", 0),
           "equatorial": (30.0, np.pi / 2, 1e-8, "30deg cap centred on the equator (interior)", 0),
           "cap15at60": (15.0, np.deg2rad(60.0), 1e-8, "15deg cap centred at colatitude 60deg (interior)", 0),
           "spin": (20.0, 0.0, 1e-8, "spin-2 field on a 20deg polar cap", 2),
           "twocaps": (15.0, 0.0, 1e-8, "two antipodal 15deg caps (multi-box leaf)", 0)}
    deg, tc, tol, desc, spin = geo[args.config]
    M = int(round(4 * (L * np.deg2rad(deg)) ** 2))
    rng = np.random.default_rng(30)
    th, ph = centered_cap(rng, deg, M, tc, 0.5)
    method = "stretch"
    if args.config == "twocaps":
        th2, ph2 = centered_cap(rng, deg, M, np.pi, 2.0)
        th, ph = np.concatenate([th, th2]), np.concatenate([ph, ph2])
        M = 2 * M
        method = "auto"
    alm = rand_alm(L, rng)
    f = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    mixed = args.precision == "mixed"
    if mixed and tol < 1e-6:
        tol = 1e-6  # the tightest tolerance mixed precision serves
    log(f"{args.config}: plan lmax={L} M={M} tol={tol} precision={args.precision}")
    t0 = time.perf_counter()
    plan = pfsht.Plan(th, ph, lmax=L, tol=tol, method=method, precision=args.precision, spin=spin)
    torch.cuda.synchronize()
    t_plan = time.perf_counter() - t0
    info = plan.info
    log(f"plan {t_plan:.2f}s kind={info['kind_name']} nth={info['nth']} nph={info['nph']} "
        f"ring_work={info['ring_work']} fine={info['es_n1']}x{info['es_n2']} w={info['es_w']} "
        f"nboxes={info['nboxes']} spin={info['spin']}")
    a_d, f_d = pfsht.to_device(alm), pfsht.to_device(f)
    fo = torch.empty(M, dtype=torch.complex128, device="cuda")
    ao = torch.empty(alm.size, dtype=torch.complex128, device="cuda")
    s = torch.cuda.current_stream()
    for _ in range(args.warmup):
        plan.synth(fo, a_d, s)
        plan.adjoint(ao, f_d, s)
    torch.cuda.synchronize()
    steps = max(1, args.steps // 2)
    plan.set_timing(True)
    plan.stage_times("synth", reset=True)
    plan.stage_times("adjoint", reset=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        plan.synth(fo, a_d, s)
        plan.adjoint(ao, f_d, s)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    plan.set_timing(False)
    ss, sa = plan.stage_times("synth"), plan.stage_times("adjoint")
    # error against the device Direct leaf on 1e4 points
    sub = np.random.default_rng(31).choice(M, 10_000, replace=False)
    pdir = pfsht.Plan(th[sub], ph[sub], lmax=L, method="direct", spin=spin)
    fsyn = pfsht.from_device(fo, (M,))
    fdir = pdir @ alm
    err_s = float(np.abs(fsyn[sub] - fdir).max() / np.abs(fdir).max())
    fsub = np.random.default_rng(32).standard_normal(10_000) + 1j * np.random.default_rng(33).standard_normal(10_000)
    fz = np.zeros(M, complex)
    fz[sub] = fsub
    ast_ = plan.H @ fz
    adir = pdir.H @ fsub
    err_a = float(np.abs(ast_ - adir).max() / np.abs(adir).max())
    del pdir
    log(f"{steps} steps: {el / steps * 1e3:.3f} ms/step; error vs Direct: synthesis {err_s:.2e}, adjoint {err_a:.2e}")
    W = info["ring_work"]
    # the adjoint ring pass: its stage time (none for a multi-box plan, whose boxes are timed by
    # their own plans); 12 flops per scalar triple (one triple serves orders +-m), 9 per spin triple
    # (recurrence 5, complex accumulate 4; orders +-m are separate triples)
    ring_ms = sa["ring"][0] / max(sa["ring"][1], 1) + (sa["ring_reduce"][0] / max(sa["ring_reduce"][1], 1) if spin else 0)
    fl = 9.0 if spin else 12.0
    tf = fl * W / (ring_ms * 1e-3) / 1e12 if ring_ms > 0 else None
    kname = "spin_adj_kernel (spin adjoint ring pass)" if spin else "ring_adj_kernel (adjoint ring pass)"
    roof = None if tf is None else {"kernel": kname, "bound": "valu", "achieved": tf, "peak": PEAK_FP64_TFLOPS,
                                    "unit": "TFLOP/s", "frac": tf / PEAK_FP64_TFLOPS, "traffic": None,
                                    "algorithmic": f"{fl:g} FP64 flops per retained triple x {W} triples"}
    out = {"metric": f"applies/s at lmax=2048 tol={tol:g} (synth+adjoint)", "value": 2 * steps / el,
           "unit": "applies/s", "n_gpus": 1, "steps": steps, "warmup": args.warmup,
           "ms_per_step": el / steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64 recurrence, f32 grid and FFT" if mixed else "f64", "data": "synthetic",
           "config": {"workload": f"{desc}, lmax={L}, M={M} area-uniform, tol={tol:g}, ntrans=1, "
                                  f"{args.precision}, one synthesis + one adjoint per step",
                      "precision": args.precision, "spin": spin, "nboxes": info["nboxes"],
                      "lmax": L, "npts": M, "tol": tol, "kind": info["kind_name"], "nth": info["nth"],
                      "nph": info["nph"], "ring_work": W, "fine_grid": [info["es_n1"], info["es_n2"]], "fine_rows_stored": info["es_rows"],
                      "es_w": info["es_w"]},
           "error_vs_direct_1e4": {"synthesis": err_s, "adjoint": err_a},
           "plan_s": t_plan,
           "stage_ms": {"synth": {k: v[0] / max(v[1], 1) for k, v in ss.items()},
                        "adjoint": {k: v[0] / max(v[1], 1) for k, v in sa.items()}},
           "roofline": roof,
           "cpu_baseline": None}
    print(json.dumps(out))


def serve_leg(ngeo, nreq, seed=555, budget=200 * 10**9, load=0.7):
    """Serving: ngeo recurring geometries, each a cap of radius U(5, 30) deg around a random centre on
    the sphere, lmax from {256, 512, 1024, 2048}, M = 4 (L theta*)^2 area-uniform points, ntrans
    U{1..8}, tol from {1e-8, 1e-5}, method chosen by the planner (Direct where the stretch is refused
    or dearer; tol >= 1e-6 plans keep FP32 grids). Requests draw a geometry and a direction (50/50).
    Closed loop: all nreq requests submitted at once -> capacity. Open loop: Poisson arrivals at
    load x capacity -> latency = completion (device clock) - arrival. One at a time: the same
    requests' transforms applied singly, in order, on ntrans = 1 plans of the same geometries and
    precision, one stream. Applies count transforms (a request applies ntrans)."""
    import torch
    import pfsht
    rng = np.random.default_rng(seed)
    geos = []
    for g in range(ngeo):
        deg = rng.uniform(5.0, 30.0)
        tc, pc = np.arccos(rng.uniform(-1.0, 1.0)), rng.uniform(0.0, 2 * np.pi)
        L = int(rng.choice([256, 512, 1024, 2048]))
        M = int(round(4 * (L * np.deg2rad(deg)) ** 2))
        B = int(rng.integers(1, 9))
        tol = float(rng.choice([1e-8, 1e-5]))
        th, ph = centered_cap(rng, deg, M, tc, pc)
        geos.append(dict(deg=deg, tc=tc, L=L, M=M, B=B, tol=tol, th=th, ph=ph))
    srv = pfsht.Server(nstreams=4, hbm_budget_bytes=budget)
    t0 = time.perf_counter()
    methods = []
    for g, G in enumerate(geos):
        info, _ = srv.plan(g, G["th"], G["ph"], G["L"], G["tol"], G["B"], "auto")
        methods.append("direct" if info["method"] == 1 else "stretch")
        LB = (G["L"] + 1) * (2 * G["L"] + 1)
        G["alm"] = torch.randn(LB * G["B"], dtype=torch.complex128, device="cuda")
        G["f"] = torch.randn(G["M"] * G["B"], dtype=torch.complex128, device="cuda")
        G["alm_out"] = torch.empty_like(G["alm"])
        G["f_out"] = torch.empty_like(G["f"])
    torch.cuda.synchronize()
    t_plans = time.perf_counter() - t0
    log(f"serve: {ngeo} plans built in {t_plans:.1f}s ({methods.count('direct')} Direct); stats {srv.stats()}")

    def submit(g, d):
        G = geos[g]
        if d == 0:
            return srv.submit(g, "synth", G["alm"], G["f_out"])
        return srv.submit(g, "adjoint", G["f"], G["alm_out"])

    for g in range(ngeo):  # warm every plan in both directions
        srv.wait(submit(g, 0))
        srv.wait(submit(g, 1))
    req = [(int(rng.integers(0, ngeo)), int(rng.integers(0, 2))) for _ in range(nreq)]
    applied = sum(geos[g]["B"] for g, _ in req)
    # closed loop: every request in flight at once -> capacity
    t0 = time.perf_counter()
    for t in [submit(g, d) for g, d in req]:
        srv.wait(t)
    t_closed = time.perf_counter() - t0
    cap_rps = nreq / t_closed
    lam = load * cap_rps
    log(f"serve: closed loop {nreq} requests in {t_closed * 1e3:.1f} ms ({applied / t_closed:.0f} applies/s); "
        f"offered load {lam:.1f} requests/s")
    arrive = np.cumsum(rng.exponential(1.0 / lam, nreq))
    # open loop: each request submitted at its arrival time; completions from the device clock
    srv.epoch()
    start = time.perf_counter()
    tickets = []
    for i in range(nreq):
        while time.perf_counter() - start < arrive[i]:
            pass
        g, d = req[i]
        tickets.append(submit(g, d))
    ends = np.array([srv.times(t)[1] for t in tickets]) * 1e-3
    el = float(ends.max())
    lat_ms = (ends - arrive) * 1e3
    st = srv.stats()
    srv.close()
    del srv
    # one at a time on ntrans = 1 plans of the same geometries and precision
    one = []
    for G in geos:
        prec = "mixed" if G["tol"] >= 1e-6 else "fp64"
        p1 = pfsht.Plan(G["th"], G["ph"], lmax=G["L"], tol=G["tol"], ntrans=1, method="auto", precision=prec)
        LB = (G["L"] + 1) * (2 * G["L"] + 1)
        one.append((p1, LB))
    s = torch.cuda.current_stream()
    for g, G in enumerate(geos):
        p1, LB = one[g]
        p1.synth(G["f_out"][:G["M"]], G["alm"][:LB], s)
        p1.adjoint(G["alm_out"][:LB], G["f"][:G["M"]], s)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for g, d in req:
        G = geos[g]
        p1, LB = one[g]
        M = G["M"]
        for b in range(G["B"]):
            if d == 0:
                p1.synth(G["f_out"][b * M:(b + 1) * M], G["alm"][b * LB:(b + 1) * LB], s)
            else:
                p1.adjoint(G["alm_out"][b * LB:(b + 1) * LB], G["f"][b * M:(b + 1) * M], s)
    torch.cuda.synchronize()
    t_one = time.perf_counter() - t0
    del one
    served = applied / t_closed
    log(f"serve: one at a time {applied / t_one:.0f} applies/s; served / one at a time {served * t_one / applied:.2f}; "
        f"open-loop p50 {np.percentile(lat_ms, 50):.2f} ms p99 {np.percentile(lat_ms, 99):.2f} ms")
    return {"geometries": ngeo, "requests": nreq, "transforms": applied,
            "direct_plans": methods.count("direct"), "plan_build_s": t_plans,
            "applies_s": served, "one_at_a_time_applies_s": applied / t_one,
            "served_over_one_at_a_time": served * t_one / applied,
            "capacity_rps": cap_rps, "offered_rps": lam, "open_loop_applies_s": applied / el,
            "latency_ms": {"p50": float(np.percentile(lat_ms, 50)), "p90": float(np.percentile(lat_ms, 90)),
                           "p99": float(np.percentile(lat_ms, 99)), "max": float(lat_ms.max())},
            "server_stats": st}


def run_serve(args):
    """The serving workload at 32 geometries (see serve_leg)."""
    import torch
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    nreq = max(200, 40 * args.steps)
    r = serve_leg(32, nreq)
    out = {"metric": "served p50/p99 latency and applies/s (mixed lmax 256-2048, tol 1e-8/1e-5, ntrans 1-8)",
           "value": r["applies_s"], "unit": "applies/s", "n_gpus": 1, "steps": nreq, "warmup": 64,
           "ms_per_step": 1e3 * r["transforms"] / r["applies_s"] / nreq, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "serving: 32 recurring geometries (cap radius U(5,30) deg, random centre, "
                                  "lmax in {256,512,1024,2048}, M = 4 (L theta*)^2, ntrans U{1..8}, tol in "
                                  "{1e-8,1e-5}, synthesis/adjoint 50/50), 4 streams, 200 GB HBM budget; value "
                                  "= closed-loop transforms/s, latency from Poisson arrivals at 70% of capacity"},
           "served": r, "roofline": None, "cpu_baseline": None}
    print(json.dumps(out))


def run_smallplans(args):
    """64 small plans (lmax 256, caps of radius U(5, 20) deg around random centres, M = 4 (L theta*)^2
    area-uniform points, tol 1e-8, the planner's method, ntrans 1): one synthesis and one adjoint of
    every plan per step, applied one at a time on one stream against the same applies as one plan
    group (pfsht_group_*: the ring passes of all 64 in one launch over a (plan, order, ring tile) work
    list, the other stages on concurrent streams, replayed from hipGraphs). value = applies/s of the
    group."""
    import torch
    import pfsht
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    L, n = 256, 64
    rng = np.random.default_rng(64)
    plans, alms, fs = [], [], []
    t0 = time.perf_counter()
    for _ in range(n):
        deg = rng.uniform(5.0, 20.0)
        tc, pc = np.arccos(rng.uniform(-1.0, 1.0)), rng.uniform(0.0, 2 * np.pi)
        M = int(round(4 * (L * np.deg2rad(deg)) ** 2))
        th, ph = centered_cap(rng, deg, M, tc, pc)
        plans.append(pfsht.Plan(th, ph, lmax=L, tol=1e-8, method="auto"))
        alms.append(pfsht.to_device(rand_alm(L, rng)))
        fs.append(torch.randn(M, dtype=torch.complex128, device="cuda"))
    torch.cuda.synchronize()
    t_plans = time.perf_counter() - t0
    fo = [torch.empty_like(f) for f in fs]
    ao = [torch.empty_like(a) for a in alms]
    g = pfsht.Group(plans)
    s = torch.cuda.current_stream()
    log(f"{n} plans in {t_plans:.1f}s ({sum(p.algorithm == 'direct' for p in plans)} Direct), "
        f"{sum(p.npts for p in plans)} points")

    def one_at_a_time():
        for p, a, f, x, y in zip(plans, alms, fs, fo, ao):
            p.synth(x, a, s)
            p.adjoint(y, f, s)

    def grouped():
        g.synth(fo, alms, s)
        g.adjoint(ao, fs, s)

    for _ in range(args.warmup):
        one_at_a_time()
        grouped()
    torch.cuda.synchronize()
    el = {"one_at_a_time": 0.0, "group": 0.0}
    rounds = 4
    per = max(1, args.steps // rounds)
    for _ in range(rounds):  # alternating rounds
        for name, fn in (("one_at_a_time", one_at_a_time), ("group", grouped)):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(per):
                fn()
            torch.cuda.synchronize()
            el[name] += time.perf_counter() - t0
    steps = rounds * per
    rate = {k: 2 * n * steps / v for k, v in el.items()}
    log(f"one at a time {rate['one_at_a_time']:.0f} applies/s, group {rate['group']:.0f} applies/s, ratio "
        f"{rate['group'] / rate['one_at_a_time']:.2f}")
    out = {"metric": "applies/s over 64 small plans (lmax 256, caps 5-20 deg, tol 1e-8), synth+adjoint",
           "value": rate["group"], "unit": "applies/s", "n_gpus": 1, "steps": steps, "warmup": args.warmup,
           "ms_per_step": el["group"] / steps * 1e3, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "64 small plans: lmax 256, cap radius U(5,20) deg, random centres, M = 4 (L theta*)^2, "
                                  "tol 1e-8, ntrans 1, one synthesis + one adjoint of each per step",
                      "plans": n, "direct_plans": sum(p.algorithm == "direct" for p in plans),
                      "points": sum(p.npts for p in plans)},
           "one_at_a_time_applies_s": rate["one_at_a_time"],
           "group_over_one_at_a_time": rate["group"] / rate["one_at_a_time"],
           "one_at_a_time_ms_per_step": el["one_at_a_time"] / steps * 1e3,
           "roofline": None, "cpu_baseline": None}
    print(json.dumps(out))


def run_tune(args):
    """The plan-time tuner at lmax 256, 1024 and 2048 (20 deg polar caps, M = 4 (L theta*)^2 up to 2e6,
    tol 1e-8, ntrans 1; and lmax 1024 at ntrans 16): the candidates each plan timed, the shape it
    chose, and a synthesis + adjoint step with the chosen shape against the default shape."""
    import torch
    import pfsht
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    rows = []
    for L, B in ((256, 1), (1024, 1), (2048, 1), (1024, 16)):
        M = min(int(round(4 * (L * np.deg2rad(20.0)) ** 2)), 2_000_000)
        rng = np.random.default_rng(L + B)
        th, ph = cap_points(rng, 20, M)
        pfsht.tune_clear()
        tuned = pfsht.Plan(th, ph, lmax=L, tol=1e-8, ntrans=B, method="stretch", tune=1)
        table = tuned.tune_table()
        pfsht.tune_clear()
        default = pfsht.Plan(th, ph, lmax=L, tol=1e-8, ntrans=B, method="stretch", tune=-1)
        a = pfsht.to_device(np.asfortranarray(np.stack([rand_alm(L, rng) for _ in range(B)], axis=2)))
        f = torch.randn(M * B, dtype=torch.complex128, device="cuda")
        fo = torch.empty_like(f)
        ao = torch.empty_like(a)
        s = torch.cuda.current_stream()
        plans = (("tuned", tuned), ("default", default))
        for _, p in plans:
            for _ in range(args.warmup):
                p.synth(fo, a, s)
                p.adjoint(ao, f, s)
            p.set_timing(True)
            p.stage_times("synth", reset=True)
            p.stage_times("adjoint", reset=True)
        # the two plans in alternating rounds, so that clock and cache state do not favour one
        el = {name: 0.0 for name, _ in plans}
        rounds = 4
        for _ in range(rounds):
            for name, p in plans:
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                for _ in range(max(1, args.steps // rounds)):
                    p.synth(fo, a, s)
                    p.adjoint(ao, f, s)
                torch.cuda.synchronize()
                el[name] += time.perf_counter() - t0
        nst = rounds * max(1, args.steps // rounds)
        res = {}
        for name, p in plans:
            p.set_timing(False)
            ss, sa = p.stage_times("synth"), p.stage_times("adjoint")
            ring = ss["ring"][0] / max(ss["ring"][1], 1) + sa["ring"][0] / max(sa["ring"][1], 1) + \
                sa["ring_reduce"][0] / max(sa["ring_reduce"][1], 1)
            res[name] = {"shape": [p.info["seg_blocks"], p.info["adj_tiles"], p.info["mfma_tiles"]],
                         "ms_per_step": el[name] / nst * 1e3, "ring_ms_per_step": ring}
        row = {"lmax": L, "ntrans": B, "npts": M, "nrings": tuned.info["nth_half"],
               "candidates": table, **res, "ring_speedup": res["default"]["ring_ms_per_step"] /
               res["tuned"]["ring_ms_per_step"]}
        log(f"lmax {L} ntrans {B}: shape {res['tuned']['shape']} ring {res['tuned']['ring_ms_per_step']:.4f} ms vs "
            f"default {res['default']['shape']} {res['default']['ring_ms_per_step']:.4f} ms; step "
            f"{res['tuned']['ms_per_step']:.4f} vs {res['default']['ms_per_step']:.4f} ms")
        rows.append(row)
        del tuned, default
    out = {"metric": "ring-pass launch shape chosen by the plan-time tuner", "value": None, "unit": None,
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "higher_is_better": True,
           "dtype": "f64", "data": "synthetic",
           "config": {"workload": "20deg polar caps, tol 1e-8, shapes timed at plan time (tune = 1)"},
           "rows": rows, "roofline": None, "cpu_baseline": None}
    print(json.dumps(out))


def run_batched(args):
    """ntrans = 16 independent coefficient sets on one plan (lmax 1024, M = 5.1e5, 20 deg cap,
    tol 1e-8, C_l = (l+1)^-2 spectra with distinct seeds), against the same applies made one at a time
    on an ntrans = 1 plan of the same points. value counts single-transform applies per second."""
    import torch
    import pfsht
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    L, tol, M, B = 1024, 1e-8, 510_000, 16
    rng = np.random.default_rng(1024)
    th, ph = cap_points(rng, 20, M)
    alm = np.asfortranarray(np.stack([rand_alm(L, np.random.default_rng(5000 + b)) for b in range(B)], axis=2))
    f = np.asfortranarray(rng.standard_normal((M, B)) + 1j * rng.standard_normal((M, B)))
    log(f"plans lmax={L} M={M} ntrans={B} and 1")
    pB = pfsht.Plan(th, ph, lmax=L, tol=tol, ntrans=B, method="stretch")
    p1 = pfsht.Plan(th, ph, lmax=L, tol=tol, ntrans=1, method="stretch")
    torch.cuda.synchronize()
    W = pB.info["ring_work"]
    a_d = pfsht.to_device(alm)
    f_d = pfsht.to_device(f)
    fo = torch.empty(M * B, dtype=torch.complex128, device="cuda")
    ao = torch.empty(alm.size, dtype=torch.complex128, device="cuda")
    LB = (L + 1) * (2 * L + 1)
    a_cols = [a_d[b * LB:(b + 1) * LB] for b in range(B)]
    f_cols = [f_d[b * M:(b + 1) * M] for b in range(B)]
    fo1 = torch.empty(M, dtype=torch.complex128, device="cuda")
    ao1 = torch.empty(LB, dtype=torch.complex128, device="cuda")
    s = torch.cuda.current_stream()
    for _ in range(args.warmup):
        pB.synth(fo, a_d, s)
        pB.adjoint(ao, f_d, s)
        p1.synth(fo1, a_cols[0], s)
        p1.adjoint(ao1, f_cols[0], s)
    torch.cuda.synchronize()
    steps = max(1, args.steps // 4)
    # both legs are timed without stage events; the stage table comes from a separate pass
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        pB.synth(fo, a_d, s)
        pB.adjoint(ao, f_d, s)
    torch.cuda.synchronize()
    tB = time.perf_counter() - t0
    pB.set_timing(True)
    pB.stage_times("synth", reset=True)
    pB.stage_times("adjoint", reset=True)
    for _ in range(steps):
        pB.synth(fo, a_d, s)
        pB.adjoint(ao, f_d, s)
    torch.cuda.synchronize()
    pB.set_timing(False)
    ss, sa = pB.stage_times("synth"), pB.stage_times("adjoint")
    t0 = time.perf_counter()
    for _ in range(steps):
        for b in range(B):
            p1.synth(fo1, a_cols[b], s)
            p1.adjoint(ao1, f_cols[b], s)
    torch.cuda.synchronize()
    t1 = time.perf_counter() - t0
    batched = 2 * B * steps / tB
    single = 2 * B * steps / t1
    log(f"batched {batched:.1f} applies/s, one at a time {single:.1f} applies/s, ratio {batched / single:.2f}")
    ring_ms = sa["ring"][0] / max(sa["ring"][1], 1)
    flops = (4.0 + 8.0 * B) * W
    tf = flops / (ring_ms * 1e-3) / 1e12
    out = {"metric": "applies/s at lmax=1024 tol=1e-8, ntrans=16 (synth+adjoint)", "value": batched,
           "unit": "applies/s", "n_gpus": 1, "steps": steps, "warmup": args.warmup,
           "ms_per_step": tB / steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"batched apply, 20deg polar cap, lmax={L}, M={M} area-uniform, tol={tol}, "
                                  f"ntrans={B}, one synthesis + one adjoint of all {B} transforms per step",
                      "lmax": L, "npts": M, "ntrans": B, "ring_work_per_transform": W},
           "one_at_a_time_applies_s": single, "batched_over_single": batched / single,
           "stage_ms": {"synth": {k: v[0] / max(v[1], 1) for k, v in ss.items()},
                        "adjoint": {k: v[0] / max(v[1], 1) for k, v in sa.items()}},
           "roofline": {"kernel": f"ring_adj_mfma_kernel (adjoint ring pass on FP64 MFMA, {B} transforms)",
                        "bound": "mfma",
                        "achieved": tf, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                        "frac": tf / PEAK_FP64_TFLOPS, "traffic": None,
                        "algorithmic": f"(4 + 8 x {B}) FP64 flops per retained triple x {W} triples"},
           "cpu_baseline": None}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
