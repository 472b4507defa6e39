"""Coefficients of the polynomial G(x) ~ acos(x) / sqrt(1 - x^2) on x in
[0, 1] (G(1) = 1) used by chk_ltc_dev.h's edge term theta / sin(theta) with
x = cos(theta): mpmath's Chebyshev fit at 40 digits, checked in FP64 Horner
evaluation.  Run: python gen_edge_poly.py [degree]."""
import sys

import mpmath as mp
import numpy as np

mp.mp.dps = 40


def g(x):
    x = mp.mpf(x)
    if x == 1:
        return mp.mpf(1)
    return mp.acos(x) / mp.sqrt(1 - x * x)


def main():
    deg = int(sys.argv[1]) if len(sys.argv) > 1 else 20
    coeffs, err = mp.chebyfit(g, [0, 1], deg + 1, error=True)
    c = [float(v) for v in coeffs]  # highest power first
    xs = np.linspace(0, 1, 20001)
    p = np.zeros_like(xs)
    for v in c:
        p = p * xs + v
    exact = np.array([float(g(v)) for v in xs])
    rel = np.abs(p - exact) / exact
    print(f"// degree {deg}: fit error {float(err):.2e}, FP64 Horner max rel error {rel.max():.2e}")
    print(",\n".join(f"    {v!r}" for v in c))


if __name__ == "__main__":
    main()
