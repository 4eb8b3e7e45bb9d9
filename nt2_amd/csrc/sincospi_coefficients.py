"""Fit the minimax polynomials of sin(pi f) and cos(pi f), |f| <= 1/2, used by
the FAST column fold (colfold.hip: sinpi_poly, cospi_poly).

sin(pi f) = f (c1 + c3 z + c5 z^2 + c7 z^3 + c9 z^4),   z = f^2
cos(pi f) = 1 + z (d2 + d4 z + d6 z^2 + d8 z^3 + d10 z^4)

Lawson's iteration (iteratively reweighted least squares, converging to the
minimax solution) on Chebyshev nodes of [0, 1/2] against mpmath's sin and
cos at 100 bits; the coefficients are then rounded to float.  Prints them as
C hexadecimal float literals and the largest absolute error of the rounded
polynomials (evaluated in double) over a dense grid.

    python3 nt2_amd/csrc/sincospi_coefficients.py
"""
from __future__ import annotations

import numpy as np


def _truth(f, fn):
    import mpmath as mp

    mp.mp.prec = 100
    g = mp.sin if fn == "sin" else mp.cos
    return np.array([float(g(mp.pi * mp.mpf(float(x)))) for x in f])


def _lawson(A, b, iters=400):
    w = np.full(len(b), 1.0 / len(b))
    c = None
    for _ in range(iters):
        W = np.sqrt(w)
        c, *_ = np.linalg.lstsq(A * W[:, None], b * W, rcond=None)
        r = np.abs(A @ c - b)
        w = w * r
        w /= w.sum()
    return c


def fit(nodes: int = 4000):
    f = np.cos(np.linspace(0.0, np.pi, nodes)) * 0.25 + 0.25
    f = np.sort(np.concatenate([f, [0.5, 1e-4]]))
    z = f * f
    sin_c = _lawson(np.stack([f * z ** k for k in range(5)], 1), _truth(f, "sin"))
    cos_c = _lawson(np.stack([z ** k for k in range(1, 6)], 1), _truth(f, "cos") - 1.0)
    return sin_c.astype(np.float32), cos_c.astype(np.float32)


def eval_sin(c, f):
    z = f * f
    p = np.float64(c[-1])
    for k in range(len(c) - 2, -1, -1):
        p = p * z + np.float64(c[k])
    return f * p


def eval_cos(c, f):
    z = f * f
    p = np.float64(c[-1])
    for k in range(len(c) - 2, -1, -1):
        p = p * z + np.float64(c[k])
    return 1.0 + z * p


def c_literal(x) -> str:
    return float(np.float32(x)).hex().replace("0000000p", "p") + "f"


if __name__ == "__main__":
    s, c = fit()
    grid = np.linspace(0.0, 0.5, 4001)
    es = np.max(np.abs(eval_sin(s, grid) - _truth(grid, "sin")))
    ec = np.max(np.abs(eval_cos(c, grid) - _truth(grid, "cos")))
    print("sin(pi f): c1..c9 =", ", ".join(c_literal(x) for x in s), f"  max |err| = {es / 2**-24:.3f} x 2^-24")
    print("cos(pi f): d2..d10 =", ", ".join(c_literal(x) for x in c), f"  max |err| = {ec / 2**-24:.3f} x 2^-24")
