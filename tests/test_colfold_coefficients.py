"""The sin(pi f) / cos(pi f) polynomials of the FAST column fold: the
literals in colfold.hip are what the committed fit reproduces, and the
rounded polynomials stay within 1 x 2^-24 of sin and cos on |f| <= 1/2."""
import os
import re
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "nt2_amd", "csrc"))
import sincospi_coefficients as sc  # noqa: E402


def _literals(fn):
    src = open(os.path.join(ROOT, "nt2_amd", "csrc", "colfold.hip")).read()
    body = src[src.index(fn):]
    body = body[: body.index("\n}")]
    return [float.fromhex(h) for h in re.findall(r"(-?0x1\.[0-9a-f]+p[+-]\d+)f", body)]


def test_literals_are_the_fit():
    s, c = sc.fit()
    sin_lit = _literals("float sinpi_poly(")  # Horner order: c9, c7, c5, c3, c1
    cos_lit = _literals("float cospi_poly(")  # d10, d8, d6, d4, d2
    assert sin_lit == [float(x) for x in s[::-1]]
    assert cos_lit == [float(x) for x in c[::-1]]


def test_polynomials_within_one_ulp_of_one():
    s = np.array(_literals("float sinpi_poly(")[::-1], np.float32)
    c = np.array(_literals("float cospi_poly(")[::-1], np.float32)
    f = np.linspace(-0.5, 0.5, 200001)
    assert np.max(np.abs(sc.eval_sin(s, f) - np.sin(np.pi * f))) < 2.0 ** -24
    assert np.max(np.abs(sc.eval_cos(c, f) - np.cos(np.pi * f))) < 2.0 ** -24
    # a slip in the fifth digit of one of the three leading coefficients
    # breaks the bound (the higher ones are pinned by the literal check)
    for k in range(3):
        s2 = s.copy()
        s2[k] = s2[k] * np.float32(1 + 1e-5)
        c2 = c.copy()
        c2[k] = c2[k] * np.float32(1 + 1e-5)
        assert np.max(np.abs(sc.eval_sin(s2, f) - np.sin(np.pi * f))) > 2.0 ** -24
        assert np.max(np.abs(sc.eval_cos(c2, f) - np.cos(np.pi * f))) > 2.0 ** -24
