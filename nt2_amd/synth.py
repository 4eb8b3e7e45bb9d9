"""Deterministic synthetic inputs for the benchmark and parity workloads.

Counter-based SplitMix64: value i of stream ``array_id`` under ``seed`` is
mix(base(seed, array_id) + (i+1) * golden), so any slice can be generated
independently and the first values of each stream are pinned by fixtures.
Uniform variates use the top 53 bits (u in [0,1)); float outputs are formed
in double and rounded once.  Log-uniform is 10**(lo + (hi-lo) u); normal is
Box-Muller on the pair (2i, 2i+1).
"""
from __future__ import annotations

import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
_KID = np.uint64(0xD1B54A32D192ED03)
CHUNK = 1 << 24

# stream ids of the named workloads
IDS = {"x": 1, "y": 2, "S": 3, "K": 4, "T": 5, "r": 6, "sig": 7, "w": 8, "X": 9, "s": 10, "k": 11, "p": 12,
       "arg": 13, "a": 14, "b": 15}


def _mix(z: np.ndarray) -> np.ndarray:
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def raw_bits(seed: int, array_id: int, start: int, n: int) -> np.ndarray:
    with np.errstate(over="ignore"):
        base = _mix(np.array([np.uint64(seed) * GOLDEN + np.uint64(array_id) * _KID], dtype=np.uint64))[0]
        idx = np.arange(start + 1, start + n + 1, dtype=np.uint64)
        return _mix(base + idx * GOLDEN)


def unit(seed: int, array_id: int, n: int, start: int = 0) -> np.ndarray:
    """u in [0, 1) as float64, 53 random bits."""
    out = np.empty(n, np.float64)
    for o in range(0, n, CHUNK):
        c = min(CHUNK, n - o)
        out[o:o + c] = (raw_bits(seed, array_id, start + o, c) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    return out


def uniform(seed, array_id, n, lo, hi, dtype=np.float32) -> np.ndarray:
    out = np.empty(n, dtype)
    for o in range(0, n, CHUNK):
        c = min(CHUNK, n - o)
        u = unit(seed, array_id, c, o)
        out[o:o + c] = (lo + (hi - lo) * u).astype(dtype)
    return out


def loguniform(seed, array_id, n, lo_exp, hi_exp, dtype=np.float32) -> np.ndarray:
    out = np.empty(n, dtype)
    for o in range(0, n, CHUNK):
        c = min(CHUNK, n - o)
        u = unit(seed, array_id, c, o)
        out[o:o + c] = np.power(10.0, lo_exp + (hi_exp - lo_exp) * u).astype(dtype)
    return out


def normal(seed, array_id, n, mu=0.0, sigma=1.0, dtype=np.float32) -> np.ndarray:
    out = np.empty(n, dtype)
    for o in range(0, n, CHUNK):
        c = min(CHUNK, n - o)
        u = unit(seed, array_id, 2 * c, 2 * o)
        u1, u2 = u[0::2], u[1::2]
        z = np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)
        out[o:o + c] = (mu + sigma * z).astype(dtype)
    return out


def full(n, value, dtype=np.float32) -> np.ndarray:
    return np.full(n, value, dtype)


# ------------------------------------------------------------ workloads
def saxpy_inputs(n=1 << 20, seed=0):
    """saxpy / sdot: x, y ~ U[-1, 1] (alpha = 2.5)."""
    return uniform(seed, IDS["x"], n, -1.0, 1.0), uniform(seed, IDS["y"], n, -1.0, 1.0)


def bs_inputs(n, seed=0, with_w=False):
    """Black-Scholes: S~U[5,30], K~U[1,100], T~U[0.25,10], r=0.02, sigma=0.30 (full streams); w~U[-1,1]."""
    d = {
        "S": uniform(seed, IDS["S"], n, 5.0, 30.0),
        "K": uniform(seed, IDS["K"], n, 1.0, 100.0),
        "T": uniform(seed, IDS["T"], n, 0.25, 10.0),
        "r": full(n, 0.02),
        "sig": full(n, 0.30),
    }
    if with_w:
        d["w"] = uniform(seed, IDS["w"], n, -1.0, 1.0)
    return d


# precision sweep argument distributions
SWEEP = {
    "exp": ("uniform", -20.0, 20.0), "tanh": ("uniform", -20.0, 20.0), "erfc": ("uniform", -20.0, 20.0),
    "log": ("loguniform", -30.0, 30.0), "sqrt": ("loguniform", -30.0, 30.0),
    "sin": ("uniform", -1e6, 1e6), "cos": ("uniform", -1e6, 1e6),
}


def sweep_args(fn: str, n, seed=0) -> np.ndarray:
    kind, lo, hi = SWEEP[fn]
    aid = IDS["arg"] * 100 + sorted(SWEEP).index(fn)
    if kind == "uniform":
        return uniform(seed, aid, n, lo, hi)
    return loguniform(seed, aid, n, lo, hi)


def sumdot_inputs(n, seed=0):
    """sum / dot values ~ N(1, 1e-3) (ill-conditioned for naive float summation)."""
    return normal(seed, IDS["a"], n, 1.0, 1e-3), normal(seed, IDS["b"], n, 1.0, 1e-3)


def colfold_inputs(rows=16384, cols=8192, seed=0):
    """X (cols, rows) row-major == (rows, cols) column-major, X~N(0,1); s, k, p per column."""
    X = normal(seed, IDS["X"], rows * cols).reshape(cols, rows)
    s = uniform(seed, IDS["s"], cols, 0.5, 2.0)
    k = uniform(seed, IDS["k"], cols, 0.1, 10.0)
    p = uniform(seed, IDS["p"], cols, 0.0, 2.0 * np.pi)
    return X, s, k, p
