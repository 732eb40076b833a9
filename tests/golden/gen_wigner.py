"""Generate tests/golden/wigner_d.json with mpmath (no reference code runs).

    python tests/golden/gen_wigner.py

Wigner small-d values d^l_{mn}(beta) from Wigner's explicit finite sum (independent of the
three-term degree recurrence the library and the oracle use), in the convention
d^1_{1,0}(beta) = -sin(beta)/sqrt(2):

    d^l_{mn}(b) = sum_k (-1)^(m-n+k) sqrt((l+m)!(l-m)!(l+n)!(l-n)!)
                  / ((l+n-k)! k! (m-n+k)! (l-m-k)!) cos(b/2)^(2l+n-m-2k) sin(b/2)^(m-n+2k)

evaluated with enough digits to absorb the sum's cancellation (about 0.62 l + 60). The tuples are
the ones the reference's own tests compare against BigFloat (test/test_spin.jl:23-34), plus a few
more orders and spins.
"""
import json
import os

import mpmath as mp

HERE = os.path.dirname(os.path.abspath(__file__))


def wigner_sum(l, m, n, beta):
    mp.mp.dps = int(0.62 * l) + 60
    b = mp.mpf(beta)
    c, s = mp.cos(b / 2), mp.sin(b / 2)
    f = mp.factorial
    pre = mp.sqrt(f(l + m) * f(l - m) * f(l + n) * f(l - n))
    tot = mp.mpf(0)
    for k in range(max(0, n - m), min(l + n, l - m) + 1):
        den = f(l + n - k) * f(k) * f(m - n + k) * f(l - m - k)
        tot += (-1) ** (m - n + k) * pre / den * c ** (2 * l + n - m - 2 * k) * s ** (m - n + 2 * k)
    return tot


CASES = [(512, 100, -2, 0.8), (2048, 700, -1, 0.5), (4096, 30, -2, 1.1), (1024, 0, 0, 2.0),
         (1, 1, 0, 0.37), (7, 3, -2, 0.9), (40, 11, 5, 0.9), (200, 60, -7, 0.9), (256, -200, 2, 0.3),
         (300, 250, -3, 2.8), (128, 127, 1, 1.5), (64, -64, -2, 0.05)]


def main():
    out = []
    for l, m, n, beta in CASES:
        v = wigner_sum(l, m, n, beta)
        out.append({"l": l, "m": m, "n": n, "beta": beta, "d": mp.nstr(v, 25)})
    with open(os.path.join(HERE, "wigner_d.json"), "w") as fh:
        json.dump(out, fh, indent=1)


if __name__ == "__main__":
    main()
