"""Generate the committed golden fixtures in tests/golden/ with mpmath (no reference code runs).

    python tests/golden/gen_golden.py

alf.json  — normalised associated Legendre values Pbar_l^m(cos theta) (Condon-Shortley phase), at the
            (l, m, theta) tuples the reference's own tests check against BigFloat
            (test/test_legendre.jl:7-26, examples/02:62-73). Values come from mpmath's Ferrers function
            `legenp` (hypergeometric evaluation, independent of the three-term recurrence) at 60 digits:
            Pbar = sqrt((2l+1)/(4 pi) (l-m)!/(l+m)!) P_l^m. The turning-point tuples use the reference's
            own recipe: theta_max of the 20 deg cap geometry and m = polar_mmax(theta_max, L, 1e-8),
            with the geometry/polar_mmax formulas (Windows.jl:84-141, Legendre.jl:48-57) restated here
            in mpmath.
es_cdf.json — int_{-pi/2}^{u} exp(beta (cos t - 1)) cos t dt (Windows.jl:200-207) by mpmath.quad.
"""
import json
import math
import os

import mpmath as mp

HERE = os.path.dirname(os.path.abspath(__file__))
mp.mp.dps = 60


def pbar(l, m, theta):
    am = abs(m)
    if am > l:
        return mp.mpf(0)
    x = mp.cos(mp.mpf(theta))
    norm = mp.sqrt((2 * l + 1) / (4 * mp.pi) * mp.factorial(l - am) / mp.factorial(l + am))
    v = norm * mp.legenp(l, am, x)
    if m < 0 and am % 2 == 1:
        v = -v
    return v


def polar_mmax(theta, L, eps):
    s = abs(math.sin(theta))
    z = (L + 0.5) * s
    lne = abs(math.log(max(eps, 2.2250738585072014e-308)))
    pad = 0.5 * (9.0 ** (1 / 3)) * (max(z, 1.0) ** (1 / 3)) * ((lne * lne) ** (1 / 3)) + 2
    return max(0, min(L, int(math.ceil(z + pad))))


def theta_max_pole(thhi, L, eps):
    beta = abs(math.log(eps))
    Lh = L * thhi
    ea = math.sqrt(2 * beta / Lh)
    return thhi * (1 + ea)


def main():
    tuples = [(4096, 1000, 0.4), (4096, 3000, 1.2), (8192, 4000, 0.9), (2048, 0, 1.5),
              (513, 513, 1.0), (100, 7, 3.0),
              (2048, 700, 0.35), (4096, 1400, 0.35), (8192, 1600, 0.20),
              (30, 7, 0.02), (30, 7, 0.11), (30, 7, 0.37), (30, 7, 0.8), (30, 7, 1.4), (30, 7, 2.7),
              (3, 2, 0.7), (7, 5, 0.7), (40, 13, 0.7), (40, -13, 0.7), (64, 20, 0.5), (256, 100, 0.8)]
    for L in (256, 1024, 2048, 4096, 8192):
        tm = theta_max_pole(20 * math.pi / 180, L, 1e-8)
        tuples.append((L, polar_mmax(tm, L, 1e-8), tm))
    out = []
    for (l, m, th) in tuples:
        v = pbar(l, m, th)
        out.append({"l": l, "m": m, "theta": th, "value": mp.nstr(v, 30, min_fixed=0, max_fixed=0),
                    "method": "mpmath.legenp"})
        print(l, m, th, mp.nstr(v, 17), flush=True)
    with open(os.path.join(HERE, "alf.json"), "w") as fh:
        json.dump(out, fh, indent=1)

    cdf = []
    for eps in (1e-8, 1e-5):
        beta = abs(math.log(eps))
        for u in (-1.0, -0.3, 0.0, 0.7, math.pi / 2):
            v = mp.quad(lambda t: mp.exp(beta * (mp.cos(t) - 1)) * mp.cos(t), [-mp.pi / 2, u])
            cdf.append({"beta": beta, "u": u, "value": mp.nstr(v, 30)})
    with open(os.path.join(HERE, "es_cdf.json"), "w") as fh:
        json.dump(cdf, fh, indent=1)


if __name__ == "__main__":
    main()
