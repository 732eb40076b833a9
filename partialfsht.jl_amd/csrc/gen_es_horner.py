#!/usr/bin/env python3
"""Writes es_horner_tables.inc: the exponential-of-semicircle kernel as piecewise polynomials.

    python partialfsht.jl_amd/csrc/gen_es_horner.py

For each kernel width w (2..16) at the shape parameter the library uses for it (beta = 2.30 w;
2.20, 2.26, 2.38 times w for w = 2, 3, 4: es_params in geometry.cpp), and a point x with base cell
b = ceil(x - w/2), put u = 2 (b - x) + w - 1 in [-1, 1). Cell t of the point sits at
z_t = (u + 1 - w + 2 t) / w and the kernel there is phi(z) = exp(beta (sqrt(1 - z^2) - 1)).

  interior cells 1 <= t <= w - 2: phi(z_t) as a polynomial in u;
  edge cells t = 0 and t = w - 1:  phi has a square-root branch point at |z| = 1, so the edge
      polynomial is in e = 2 sqrt((1 + s u) / 2) - 1 (s = +1 for t = 0, -1 for
      t = w - 1), i.e. v = sqrt(1 - |z|) rescaled to [-1, 1], in which
      phi = exp(beta (v sqrt(2 - v^2) - 1)) is analytic. By phi(-z) = phi(z) one edge table serves
      both ends, and p_{w-1-t}(u) = p_t(-u) for the interior ones.

Each polynomial is the Chebyshev interpolant at Chebyshev points of the first kind, computed in
60-digit arithmetic and converted to monomial coefficients (highest first for Horner), then rounded
to double. The degrees are the smallest (interior at most 14, edge at most 22) whose largest error
over the cells, sampled at 801 points, is within target(w) = max(2e-16, 1e-5 tol_w) of the kernel
(which peaks at 1), tol_w = 10^-(w-2) being the tightest tolerance that picks width w: the
kernel's own truncation error at that tolerance stays far above the polynomial's. The tables
are compile-time constants of the device code, so the coefficients are immediate operands and
never loaded.

EsHorner32<w> (w = 2..8, the widths a mixed-precision plan can pick, tol >= 1e-6) holds the same
polynomials for FP32 evaluation: the degrees are the smallest whose coefficients, rounded to float
and run through Horner in float arithmetic, stay within TARGET32 of the kernel on the same samples.
"""
import os
import struct

import mpmath as mp

mp.mp.dps = 60
HERE = os.path.dirname(os.path.abspath(__file__))
DEG_INT, DEG_EDGE = 14, 22  # largest degrees
TARGET32 = 3e-7  # FP32 tables: a few float roundings of the kernel's peak value 1
W32 = range(2, 9)


def target(w):
    return max(2e-16, 1e-5 * 10.0 ** (-(w - 2)))


def _horner(c, x):
    acc = 0.0
    for a in reversed(c):
        acc = acc * x + a
    return acc


def _maxerr(fn, c):
    xs = [-1 + 2 * i / 800 for i in range(801)]
    return max(abs(_horner(c, x) - float(fn(mp.mpf(x)))) for x in xs)


def _fit(fns, dmax, tgt):
    for d in range(4, dmax + 1):
        cs = [cheb_monomial(fn, d) for fn in fns]
        if d == dmax or max(_maxerr(fn, c) for fn, c in zip(fns, cs)) <= tgt:
            return cs


def _f32(x):
    return struct.unpack("f", struct.pack("f", x))[0]


def _maxerr32(fn, c):
    """Largest error of Horner in float arithmetic (each step rounded, as an FMA rounds once) with
    the coefficients rounded to float, at float-rounded arguments."""
    cf = [_f32(a) for a in c]
    worst = 0.0
    for i in range(801):
        x = _f32(-1 + 2 * i / 800)
        acc = 0.0
        for a in reversed(cf):
            acc = _f32(acc * x + a)
        worst = max(worst, abs(acc - float(fn(mp.mpf(x)))))
    return worst


def _fit32(fns, dmax):
    for d in range(3, dmax + 1):
        cs = [cheb_monomial(fn, d) for fn in fns]
        if d == dmax or max(_maxerr32(fn, c) for fn, c in zip(fns, cs)) <= TARGET32:
            return [[_f32(a) for a in c] for c in cs]


def es_beta(w):
    bo = 2.30
    if w == 2:
        bo = 2.20
    elif w == 3:
        bo = 2.26
    elif w == 4:
        bo = 2.38
    return bo * w  # the same IEEE double product as the C++ (both round once)


def cheb_monomial(fn, d):
    n = d + 1
    nodes = [mp.cos(mp.pi * (j + mp.mpf(1) / 2) / n) for j in range(n)]
    fv = [fn(u) for u in nodes]
    c = []
    for k in range(n):
        acc = mp.fsum(fv[j] * mp.cos(mp.pi * k * (j + mp.mpf(1) / 2) / n) for j in range(n))
        c.append(acc * 2 / n)
    c[0] /= 2
    # monomial coefficients of sum c_k T_k
    a = [mp.mpf(0)] * n
    t0, t1 = [mp.mpf(0)] * n, [mp.mpf(0)] * n
    t0[0] = mp.mpf(1)
    for i in range(n):
        a[i] += c[0] * t0[i]
    if d >= 1:
        t1[1] = mp.mpf(1)
        for i in range(n):
            a[i] += c[1] * t1[i]
    for k in range(2, n):
        tn = [mp.mpf(0)] * n
        for i in range(n - 1):
            tn[i + 1] += 2 * t1[i]
        for i in range(n):
            tn[i] -= t0[i]
            a[i] += c[k] * tn[i]
        t0, t1 = t1, tn
    return [float(x) for x in a]


def tables(w, single=False):
    beta = mp.mpf(es_beta(w))
    fns = []
    for t in range(1, w - 1):
        def fn(u, t=t):
            z = (u + 1 - w + 2 * t) / mp.mpf(w)
            zz = 1 - z * z
            return mp.e ** (beta * (mp.sqrt(zz) - 1)) if zz > 0 else mp.mpf(0)
        fns.append(fn)
    vmax = mp.sqrt(mp.mpf(2) / w)

    def edge(e):
        v = (e + 1) / 2 * vmax
        return mp.e ** (beta * (v * mp.sqrt(2 - v * v) - 1))
    if single:
        ints = _fit32(fns, DEG_INT) if fns else []
        return ints, _fit32([edge], DEG_EDGE)[0]
    ints = _fit(fns, DEG_INT, target(w)) if fns else []
    return ints, _fit([edge], DEG_EDGE, target(w))[0]


def main():
    out = ["// es_horner_tables.inc — generated by gen_es_horner.py; do not edit.",
           "// EsHorner<w>: the ES kernel of width w as piecewise polynomials (see the generator).",
           "// c: interior cells t = 1..w-2, nint monomial coefficients each (constant term first), then",
           "// the edge polynomial, nedge coefficients.",
           "template <int W> struct EsHorner;"]
    for w in range(2, 17):
        ints, edge = tables(w)
        nint = len(ints[0]) if ints else 1
        vals = [x for p in ints for x in p] + edge
        body = ",\n        ".join(", ".join(v.hex() for v in vals[i:i + 3]) for i in range(0, len(vals), 3))
        out.append(f"template <> struct EsHorner<{w}> {{\n"
                   f"    static constexpr int nint = {nint}, nedge = {len(edge)};\n"
                   f"    static constexpr double beta = {es_beta(w).hex()};\n"
                   f"    static constexpr double c[{len(vals)}] = {{\n        {body}}};\n}};")
        print(f"w {w}: interior degree {nint - 1}, edge degree {len(edge) - 1}", flush=True)
    out.append("// EsHorner32<w>: the same polynomials for FP32 evaluation (float coefficients, degrees fitted")
    out.append("// for Horner in float arithmetic), for the widths of mixed-precision plans.")
    out.append("template <int W> struct EsHorner32;")
    for w in W32:
        ints, edge = tables(w, single=True)
        nint = len(ints[0]) if ints else 1
        vals = [x for p in ints for x in p] + edge
        body = ",\n        ".join(", ".join(v.hex() + "f" for v in vals[i:i + 3]) for i in range(0, len(vals), 3))
        out.append(f"template <> struct EsHorner32<{w}> {{\n"
                   f"    static constexpr int nint = {nint}, nedge = {len(edge)};\n"
                   f"    static constexpr float c[{len(vals)}] = {{\n        {body}}};\n}};")
        print(f"w {w} (FP32): interior degree {nint - 1}, edge degree {len(edge) - 1}", flush=True)
    with open(os.path.join(HERE, "es_horner_tables.inc"), "w") as fh:
        fh.write("\n".join(out) + "\n")


if __name__ == "__main__":
    main()
