"""Coefficients of the polynomial P(w) ~ atan(sqrt(w)) / sqrt(w) on w in
[0, 1] used by chk_ltc_dev.h's atan2 (atan(t) = t P(t^2) for t in [0, 1]):
mpmath's Chebyshev fit at 40 digits, checked in FP64 Horner evaluation
against mpmath's atan.  Run: python gen_atan_poly.py [degree]."""
import sys

import mpmath as mp
import numpy as np

mp.mp.dps = 40


def f(w):
    return mp.mpf(1) if w == 0 else mp.atan(mp.sqrt(w)) / mp.sqrt(w)


def main():
    deg = int(sys.argv[1]) if len(sys.argv) > 1 else 19
    coeffs, err = mp.chebyfit(f, [0, 1], deg + 1, error=True)
    c = [float(v) for v in coeffs]  # highest power first
    t = np.linspace(0, 1, 200001)
    w = t * t
    p = np.zeros_like(w)
    for v in c:
        p = p * w + v
    approx = t * p
    exact = np.array([float(mp.atan(mp.mpf(float(x)))) for x in t[::50]])
    rel = np.abs(approx[::50] - exact) / np.maximum(np.abs(exact), 1e-300)
    print(f"// degree {deg} in w = t^2: fit error {float(err):.2e}, FP64 Horner max rel error "
          f"{rel[1:].max():.2e}")
    print(",\n".join(f"    {v!r}" for v in c))


if __name__ == "__main__":
    main()
