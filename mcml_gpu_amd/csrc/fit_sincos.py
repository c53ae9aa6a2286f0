"""Coefficients of the sin/cos polynomials in mcml_detmath.h (mcml_sincos_turns).

On |a| <= pi/4:  sin a = a + a^3 P(a^2), P of degree 2;  cos a = 1 + a^2 Q(a^2), Q of
degree 3.  Near-minimax in absolute error, by iteratively reweighted least squares (Lawson)
on Chebyshev-spaced nodes; printed as float32 literals.  Run: python fit_sincos.py
"""
import numpy as np

X = np.pi / 4


def fit(target, weight, deg, zmax, iters=30):
    z = (1 - np.cos(np.linspace(0, np.pi, 20001))) / 2 * zmax
    z[0] = 1e-12
    w = np.ones_like(z)
    for _ in range(iters):
        V = np.vstack([z ** k for k in range(deg + 1)]).T * weight(z)[:, None]
        coef, *_ = np.linalg.lstsq(V * w[:, None], target(z) * weight(z) * w, rcond=None)
        err = V @ coef - target(z) * weight(z)
        w = w * (1 + 4 * (np.abs(err) / np.abs(err).max()) ** 2)
    return coef, np.abs(err).max()


def main():
    a = np.sqrt
    ps, es = fit(lambda z: (np.sin(a(z)) - a(z)) / a(z) ** 3, lambda z: a(z) ** 3, 2, X * X)
    qc, ec = fit(lambda z: (np.cos(a(z)) - 1) / z, lambda z: z, 3, X * X)
    print("sin P:", ", ".join("%.9ef" % np.float32(c) for c in ps), " max |error| %.2e" % es)
    print("cos Q:", ", ".join("%.9ef" % np.float32(c) for c in qc), " max |error| %.2e" % ec)


if __name__ == "__main__":
    main()
