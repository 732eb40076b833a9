"""The device NUFFT's kernel tables (partialfsht.jl_amd/csrc/es_horner_tables.inc) against the
exponential-of-semicircle kernel they stand for, phi(z) = exp(beta (sqrt(1 - z^2) - 1)), evaluated
exactly as the oracle's NUFFT does. Every width the library can pick (2..16) is checked over a fine
sweep of point offsets, each of the w cells a point touches, evaluating the polynomials in double
precision by Horner exactly as the kernels do. A dropped coefficient, a wrong cell index or a
mirrored edge fails these bounds by orders of magnitude."""
import math
import os
import re

import numpy as np
import pytest

from conftest import ROOT

INC = os.path.join(ROOT, "partialfsht.jl_amd", "csrc", "es_horner_tables.inc")


def _tables():
    txt = open(INC).read()
    out = {}
    for m in re.finditer(r"struct EsHorner<(\d+)> \{\s*static constexpr int nint = (\d+), nedge = (\d+);\s*"
                         r"static constexpr double beta = (\S+);\s*"
                         r"static constexpr double c\[(\d+)\] = \{([^}]*)\}", txt):
        w, hint, hedge = int(m.group(1)), int(m.group(2)), int(m.group(3))
        vals = [float.fromhex(v.strip()) for v in m.group(6).split(",")]
        assert len(vals) == int(m.group(5)) == max(w - 2, 0) * hint + hedge
        out[w] = (float.fromhex(m.group(4)), hint, vals)
    return out


def _horner(c, x):
    acc = np.zeros_like(x)
    for a in reversed(c):
        acc = acc * x + a
    return acc


def _beta_of(w):
    bo = {2: 2.20, 3: 2.26, 4: 2.38}.get(w, 2.30)
    return bo * w


# bound on the largest absolute error (phi peaks at 1): the generator fits each width to
# max(2e-16, 1e-5 tol_w), tol_w = 10^-(w-2), on 801 samples; three times that plus rounding here
def _bound(w):
    return 3 * max(2e-16, 1e-5 * 10.0 ** (-(w - 2))) + 3e-16


@pytest.mark.parametrize("w", range(2, 17))
def test_tables_match_kernel(w):
    beta, hint, c = _tables()[w]
    assert beta == _beta_of(w)  # the shape parameter es_params chooses for this width
    x = 1000.0 + np.linspace(0.0, 1.0, 40001, endpoint=False)
    b = np.ceil(x - 0.5 * w)
    u = 2.0 * (b - x) + (w - 1)
    assert u.min() >= -1.0 and u.max() < 1.0
    worst = 0.0
    for t in range(w):
        # the kernel in 80-bit extended precision: a double-precision exp(beta (...)) is itself
        # off by ~beta * 1e-16
        z = (b.astype(np.longdouble) + t - x) * (np.longdouble(2) / w)
        exact = np.exp(np.longdouble(beta) * (np.sqrt(np.maximum(1 - z * z, 0)) - 1))
        if 1 <= t <= w - 2:
            got = _horner(c[(t - 1) * hint:t * hint], u)
        else:
            s = 1.0 if t == 0 else -1.0
            e = 2.0 * np.sqrt(0.5 * (1.0 + s * u)) - 1.0
            got = _horner(c[(w - 2) * hint:], e)
        worst = max(worst, float(np.abs(got.astype(np.longdouble) - exact).max()))
    assert worst < _bound(w), (w, worst)


def _tables32():
    txt = open(INC).read()
    out = {}
    for m in re.finditer(r"struct EsHorner32<(\d+)> \{\s*static constexpr int nint = (\d+), nedge = (\d+);\s*"
                         r"static constexpr float c\[(\d+)\] = \{([^}]*)\}", txt):
        w, hint, hedge = int(m.group(1)), int(m.group(2)), int(m.group(3))
        vals = [float.fromhex(v.strip().rstrip("f")) for v in m.group(5).split(",")]
        assert len(vals) == int(m.group(4)) == max(w - 2, 0) * hint + hedge
        assert all(np.float32(v) == v for v in vals)  # exactly representable in float
        out[w] = (hint, np.array(vals, np.float32))
    return out


def test_fp32_tables_cover_mixed_widths():
    # mixed-precision plans take tol >= 1e-6, i.e. widths up to ceil(-log10 tol) + 2 = 8
    assert sorted(_tables32()) == list(range(2, 9))


@pytest.mark.parametrize("w", range(2, 9))
def test_fp32_tables_match_kernel(w):
    # the FP32 evaluation as the kernels run it: u rounded to float, e = 2 sqrt((1 +- u) / 2) - 1 and
    # Horner in float arithmetic (each step rounded), against the kernel in extended precision; the
    # bound is a few float roundings of the peak value 1 (FP32 kernels serve tol >= 1e-6 only)
    hint, c = _tables32()[w]
    beta = _beta_of(w)
    x = 1000.0 + np.linspace(0.0, 1.0, 20001, endpoint=False)
    b = np.ceil(x - 0.5 * w)
    u = (2.0 * (b - x) + (w - 1)).astype(np.float32)
    one, half, two = np.float32(1), np.float32(0.5), np.float32(2)
    worst = 0.0
    for t in range(w):
        z = (b.astype(np.longdouble) + t - x) * (np.longdouble(2) / w)
        exact = np.exp(np.longdouble(beta) * (np.sqrt(np.maximum(1 - z * z, 0)) - 1))
        if 1 <= t <= w - 2:
            cc, arg = c[(t - 1) * hint:t * hint], u
        else:
            s = one if t == 0 else -one
            cc, arg = c[(w - 2) * hint:], two * np.sqrt(half * (one + s * u)) - one
        acc = np.zeros_like(arg)
        for a in reversed(cc):
            acc = (acc * arg + a).astype(np.float32)
        worst = max(worst, float(np.abs(acc.astype(np.longdouble) - exact).max()))
    assert worst < 1e-6, (w, worst)
