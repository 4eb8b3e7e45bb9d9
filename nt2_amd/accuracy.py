"""Stated accuracy of the FP32 Black-Scholes forward + adjoint kernel
(nt2h_bs_fwd_adj_f32, FUSED arithmetic) against the plain FP64 evaluation of
the same tree on the same inputs promoted to double, and the metrics that
check it.  The bounds hold on the S~U[5,30], K~U[1,100], T~U[0.25,10],
r=0.02, sigma=0.30, w~U[-1,1] inputs (measured maxima over 2^20 + 3 and
2^24 options, bound = about twice the larger); the parity test and the
benchmark line both evaluate `metrics` and compare them with `BOUNDS`.

Metrics (each a maximum over the options):
  price_ulp_big      |dprice| in float ULPs of the FP64 price where
                     price >= 1e-3 S (the difference S N(d1) - K e^{-rT} N(d2)
                     cancels: the ratio of its terms to the price enters)
  price_scaled       |dprice| / max(|price|, 1e-6 S)
  dS, dK, dT, dr, dsig
                     |dg| / (|w| x the greek's natural scale: 1, e^{-rT},
                     S, S T, S sqrt(T))
  total_dr, total_dsig
                     |d sum g| / sum |g_i| (the w-signed terms cancel, so
                     the error is stated against the sum of magnitudes);
                     the terms' own float errors average out over a
                     portfolio, so these two bounds are for portfolios of
                     the measured sizes: over a few options a total is as
                     accurate as its float terms (about 1e-7 of sum |g_i|)
"""
from __future__ import annotations

BOUNDS = {
    "price_ulp_big": 250.0,
    "price_scaled": 3.9e-5,
    "dS": 4.6e-7,
    "dK": 5.9e-7,
    "dT": 7.2e-8,
    "dr": 5e-7,
    "dsig": 2.9e-7,
    "total_dr": 1.8e-10,
    "total_dsig": 2.1e-10,
}


def _xp(a):
    if type(a).__module__.startswith("torch"):
        import torch

        return torch
    import numpy as np

    return np


def _floor(xp, a, v):
    return a.clamp(min=v) if xp.__name__ == "torch" else xp.maximum(a, v)


def _f32_ulp(xp, mag):
    """Spacing of float32 at |mag| (mag in float64, normal range)."""
    if xp.__name__ == "torch":
        m = mag.abs().clamp(min=2.0 ** -126)
        e = xp.floor(xp.log2(m))
        return xp.exp2(e - 23)
    m = xp.maximum(xp.abs(mag), 2.0 ** -126)
    return xp.exp2(xp.floor(xp.log2(m)) - 23)


def metrics(inputs64, outs32, outs64, totals32, totals64):
    """inputs64 = (S, K, T, r, sig, w) in float64; outs32 = the float kernel's
    (price, dS, dK, dT, dr, dsig); outs64 = the FP64 evaluation's; totals =
    (sum dr, sum dsig) of each.  numpy arrays or torch tensors."""
    S, K, T, r, sig, w = inputs64
    xp = _xp(S)
    f = (lambda a: a.double()) if xp.__name__ == "torch" else (lambda a: a.astype("float64"))
    g32 = [f(o) for o in outs32]
    aw = xp.abs(w)
    sqT = xp.sqrt(T)
    dp = xp.abs(g32[0] - outs64[0])
    m = {}
    big = xp.abs(outs64[0]) >= 1e-3 * S
    ratio = dp / _f32_ulp(xp, outs64[0])
    m["price_ulp_big"] = float(xp.max(ratio[big])) if bool(big.any()) else 0.0
    m["price_scaled"] = float(xp.max(dp / xp.maximum(xp.abs(outs64[0]), 1e-6 * S)))
    scales = [None, aw, aw * xp.exp(-r * T), aw * S, aw * S * T, aw * S * sqT]
    for i, nm in enumerate(("dS", "dK", "dT", "dr", "dsig"), start=1):
        e = xp.abs(g32[i] - outs64[i]) / _floor(xp, scales[i], 1e-30)
        m[nm] = float(xp.max(e))
    m["total_dr"] = abs(float(totals32[0]) - float(totals64[0])) / float(xp.sum(xp.abs(outs64[4])))
    m["total_dsig"] = abs(float(totals32[1]) - float(totals64[1])) / float(xp.sum(xp.abs(outs64[5])))
    return m


def check(m):
    """Names of the metrics above their bound (empty when all hold)."""
    return [k for k, b in BOUNDS.items() if not (m[k] <= b)]
