#!/usr/bin/env python3
"""Fit the polynomial used by mcml_log: log(1+f) ~= f + f^2 * P(f) on
f in [sqrt(1/2)-1, sqrt(2)-1], P of degree DEG, minimising the maximum relative error by
iteratively re-weighted least squares on a dense Chebyshev grid (a discrete Lawson /
Remez-style fit), then rounding the coefficients to binary32.

Usage: python fit_log1p.py [DEG]   (prints the coefficients, highest degree first)
"""
import sys

import numpy as np


def fit(deg, iters=60):
    a, b = np.sqrt(0.5) - 1.0, np.sqrt(2.0) - 1.0
    k = np.arange(20000)
    f = (a + b) / 2 + (b - a) / 2 * np.cos(np.pi * (k + 0.5) / len(k))
    y = np.log1p(f)
    # unknowns: c_j for f^(j+2), j = 0..deg; target (y - f) / f^2; error measured relative to y
    V = np.stack([f ** (j + 2) for j in range(deg + 1)], axis=1)
    w = np.ones_like(f)
    for _ in range(iters):
        W = w / np.abs(y)
        c, *_ = np.linalg.lstsq(V * W[:, None], (y - f) * W, rcond=None)
        err = np.abs((f + V @ c - y) / y)
        w = w * (err / err.max()) ** 0.5 + 1e-12
        w /= w.mean()
    return c, err.max()


def main():
    deg = int(sys.argv[1]) if len(sys.argv) > 1 else 10
    c, e = fit(deg)
    c32 = c.astype(np.float32)
    print("# degree %d, max relative error (float64 evaluation) %.3e" % (deg, e))
    for j in reversed(range(deg + 1)):
        print("f^%d: %.9e" % (j + 2, float(c32[j])))


if __name__ == "__main__":
    main()
