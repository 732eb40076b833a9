"""The LDS grid stage's transform plans (partialfsht.jl_amd/csrc/lds_fft.hpp), on the CPU: the header's
own planner, compiled for the host, gives the radix passes of every 5-smooth length up to 16384;
each plan must multiply out to the length with radices 2/3/4/5/8 in at most kFftMaxPass passes, and
the Stockham pass structure the device runs (butterfly i reads x[i + j n/R], twiddle w^(jk) with
k = i mod p, writes x[(i - k) R + k + j p]; i / p by the float reciprocal) must reproduce numpy's FFT,
with every index in range and each pass writing every element once. The per-thread capacity rule
(E values per thread on the workgroup size the grid stage picks) is checked for every pass."""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "partialfsht.jl_amd", "csrc", "lds_fft.hpp")

E = 8  # values per thread (kFftElems)


def smooth(limit):
    return sorted({2 ** a * 3 ** b * 5 ** c for a in range(15) for b in range(10) for c in range(7)
                   if 2 ** a * 3 ** b * 5 ** c <= limit})


@pytest.fixture(scope="module")
def plans(tmp_path_factory):
    hipcc = "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc) or shutil.which("g++") is None:
        pytest.skip("no hipcc")
    d = tmp_path_factory.mktemp("ldsfft")
    src = d / "plan.cpp"
    ns = smooth(16384)
    src.write_text(
        f'#include "{HDR}"\n#include <cstdio>\n'
        "int main() {\n"
        f"  const int ns[] = {{{', '.join(map(str, ns))}}};\n"
        '  std::printf("%d %d\\n", pfsht::kFftElems, pfsht::kFftMaxPass);\n'
        "  for (int n : ns) {\n"
        "    pfsht::FftLen f;\n"
        "    const bool ok = pfsht::fft_len_plan(n, f);\n"
        '    std::printf("%d %d", n, ok ? 1 : 0);\n'
        '    for (int s = 0; s < f.npass; ++s) std::printf(" %d", f.rad[s]);\n'
        '    std::printf("\\n");\n'
        "  }\n"
        '  pfsht::FftLen g;\n'
        '  std::printf("7 %d\\n", pfsht::fft_len_plan(7, g) ? 1 : 0);\n'
        "  return 0;\n}\n")
    exe = d / "plan"
    r = subprocess.run([hipcc, "-std=c++17", "-O1", "-o", str(exe), str(src)], capture_output=True, text=True)
    if r.returncode != 0:
        pytest.fail(r.stderr[-2000:])
    lines = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines()
    elems, maxpass = map(int, lines[0].split())
    out = {}
    for ln in lines[1:]:
        v = list(map(int, ln.split()))
        out[v[0]] = (bool(v[1]), v[2:])
    return elems, maxpass, out


def threads(n):
    t = (n + E - 1) // E
    return min(1024, max(64, (t + 63) // 64 * 64))


def stockham(x, rad, inv=False):
    n = len(x)
    tw = np.exp(-2j * np.pi * np.arange(n) / n)
    if inv:
        tw = tw.conj()
    p = 1
    for R in rad:
        nbf = n // R
        i = np.arange(nbf)
        g = ((i.astype(np.float32) + np.float32(0.5)) * (np.float32(1) / np.float32(p))).astype(np.int64)
        assert np.array_equal(g, i // p)
        k = i - g * p
        u = np.stack([x[i + j * nbf] for j in range(R)])                      # inputs
        u = u * tw[(np.arange(R)[:, None] * k[None, :] * (n // (p * R)))]      # w^(jk)
        s = 1 if inv else -1
        F = np.exp(s * 2j * np.pi * np.outer(np.arange(R), np.arange(R)) / R)  # R-point DFT
        v = F @ u
        dst = (i - k)[None, :] * R + k[None, :] + np.arange(R)[:, None] * p
        assert dst.min() >= 0 and dst.max() < n and len(np.unique(dst)) == n
        y = np.empty_like(x)
        y[dst] = v
        x = y
        p *= R
    assert p == n
    return x


def test_planner_covers_smooth_lengths(plans):
    elems, maxpass, out = plans
    assert elems == E
    for n, (ok, rad) in out.items():
        if n == 7:
            assert not ok  # another prime factor
            continue
        assert ok, n
        assert int(np.prod(rad)) == n and len(rad) <= maxpass and all(r in (2, 3, 4, 5, 8) for r in rad), (n, rad)
        # a radix-2 pass only when the power of two is 2 itself (2^4 = 4 x 4, not 8 x 2)
        assert rad.count(2) <= 1 and (2 not in rad or (n % 4 != 0)), (n, rad)
        # each pass's butterflies fit the workgroup: ceil(E / R) per thread
        if n <= E * 1024:
            nt = threads(n)
            for R in rad:
                assert -(-E // R) * nt >= n // R, (n, R, nt)


@pytest.mark.parametrize("inv", [False, True])
def test_stockham_passes_match_numpy(plans, inv):
    _, _, out = plans
    rng = np.random.default_rng(7)
    ns = [n for n in out if n != 7]
    pick = sorted(set(ns[:40]) | {720, 1440, 1920, 3840, 675, 4096, 6144, 9216, 10000, 16384}
                  | set(rng.choice(ns, 30, replace=False).tolist()))
    for n in pick:
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        want = np.fft.ifft(x) * n if inv else np.fft.fft(x)
        got = stockham(x.copy(), out[n][1], inv)
        assert np.abs(got - want).max() <= 1e-12 * max(1.0, np.abs(want).max()), n
