"""ctypes front of the oracle library (libnt2ref.so) — TEST INFRASTRUCTURE ONLY.

The oracle is a CPU restatement of NT2's scalar evaluation path (see
nt2ref.hpp for the per-function citations of the NT2 sources).  Only the
tests, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg import
this module; the product package ``nt2_amd`` never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# NT2REF_LIB selects another build of the same sources (the sanitizer build)
LIB_PATH = os.environ.get("NT2REF_LIB") or os.path.join(HERE, "libnt2ref.so")

# op ids (same numbering as NT2H_OP_* in include/nt2hip.h)
OPS = {
    "exp": 0, "log": 1, "sqrt": 2, "erf": 3, "erfc": 4, "normcdf": 5, "tanh": 6,
    "sin": 7, "cos": 8, "normpdf": 9, "fastnormcdf": 10, "exp2": 11, "exp10": 12,
}

_lib = None


def build(force: bool = False) -> str:
    """Compile libnt2ref.so with the oracle's Makefile (gcc, no contraction)."""
    if force and os.path.exists(LIB_PATH):
        os.remove(LIB_PATH)
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        i64 = C.c_int64
        L.nt2ref_map_f32.argtypes = [C.c_int, P, P, i64]
        L.nt2ref_map_f64.argtypes = [C.c_int, P, P, i64]
        L.nt2ref_saxpy_f32.argtypes = [C.c_float, P, P, i64]
        L.nt2ref_fold_sum_f32.argtypes = [P, i64]
        L.nt2ref_fold_sum_f32.restype = C.c_float
        L.nt2ref_fold_dot_f32.argtypes = [P, P, i64]
        L.nt2ref_fold_dot_f32.restype = C.c_float
        L.nt2ref_fold_sum_f64.argtypes = [P, i64]
        L.nt2ref_fold_sum_f64.restype = C.c_double
        L.nt2ref_exact_sum_f32.argtypes = [P, i64, P]
        L.nt2ref_exact_dot_f32.argtypes = [P, P, i64, P]
        L.nt2ref_bs_fwd_f32.argtypes = [P] * 7 + [i64]
        L.nt2ref_bs_fwd_f64.argtypes = [P] * 7 + [i64]
        L.nt2ref_bs_fwd_adj_f32.argtypes = [P] * 13 + [i64]
        L.nt2ref_bs_fwd_adj_f64.argtypes = [P] * 13 + [i64]
        L.nt2ref_cpu_bs_fwd_adj_f32.argtypes = [P] * 12 + [i64, i64]
        L.nt2ref_colfold_f32.argtypes = [P, i64, i64, i64, P, P, P, P]
        L.nt2ref_colfold_gc_f64.argtypes = [P, i64, i64, i64, P, P, P, P, P, P, P, P]
        L.nt2ref_ulp_stats_f32.argtypes = [P, P, i64, P]
        L.nt2ref_ulp_stats_f64.argtypes = [P, P, i64, P]
        L.nt2ref_cmp_f32.argtypes = [P, P, i64, P]
        L.nt2ref_nt2_ulpdist_f32.argtypes = [C.c_float, C.c_float]
        L.nt2ref_nt2_ulpdist_f32.restype = C.c_double
        L.nt2ref_nt2_ulpdist_f64.argtypes = [C.c_double, C.c_double]
        L.nt2ref_nt2_ulpdist_f64.restype = C.c_double
        L.nt2ref_exhaustive_hist_f32.argtypes = [C.c_int, C.c_float, C.c_float, P]
        L.nt2ref_max_threads.restype = C.c_int
        L.nt2ref_cpu_simd_bs_fwd_adj_f32.argtypes = [P] * 12 + [i64, i64]
        L.nt2ref_cpu_simd_map_f32.argtypes = [C.c_int, P, P, i64]
        L.nt2ref_cpu_simd_isa.restype = C.c_char_p
        L.nt2ref_cpu_simd_isa.argtypes = []
        for nm in ("nt2ref_fold_lmh_f32", "nt2ref_fold_lmh_f64", "nt2ref_scan_lmh_f32", "nt2ref_scan_lmh_f64"):
            getattr(L, nm).argtypes = [C.c_int, P, i64, i64, i64, P]
        _lib = L
    return _lib


def _p(a: np.ndarray) -> int:
    assert a.flags["C_CONTIGUOUS"] or a.flags["F_CONTIGUOUS"]
    return a.ctypes.data


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def map_f32(op: str, x) -> np.ndarray:
    x = _f32(x)
    y = np.empty_like(x)
    lib().nt2ref_map_f32(OPS[op], _p(x), _p(y), x.size)
    return y


def map_f64(op: str, x) -> np.ndarray:
    x = _f64(x)
    y = np.empty_like(x)
    lib().nt2ref_map_f64(OPS[op], _p(x), _p(y), x.size)
    return y


def saxpy_f32(a: float, x, y) -> np.ndarray:
    x = _f32(x)
    out = _f32(y).copy()
    lib().nt2ref_saxpy_f32(a, _p(x), _p(out), x.size)
    return out


def fold_sum_f32(x) -> np.float32:
    x = _f32(x)
    return np.float32(lib().nt2ref_fold_sum_f32(_p(x), x.size))


FOLD_OPS = {"sum": 0, "prod": 1, "maximum": 2, "minimum": 3, "asum1": 4, "asum2": 5, "norm2": 6, "mean": 7}
SCAN_OPS = {"cumsum": 0, "cumprod": 1, "cummax": 2, "cummin": 3}


def fold_lmh(op: str, x, lo: int, mid: int, hi: int) -> np.ndarray:
    """NT2's sequential fold over the middle index of a contiguous (lo, mid, hi)
    table; returns lo*hi values (index i + lo*h)."""
    f64 = np.asarray(x).dtype == np.float64
    x = _f64(x) if f64 else _f32(x)
    assert x.size == lo * mid * hi
    out = np.empty(lo * hi, x.dtype)
    fn = lib().nt2ref_fold_lmh_f64 if f64 else lib().nt2ref_fold_lmh_f32
    fn(FOLD_OPS[op], _p(x), lo, mid, hi, _p(out))
    return out


def scan_lmh(op: str, x, lo: int, mid: int, hi: int) -> np.ndarray:
    """NT2's sequential scan (scan_step) over the middle index of a (lo, mid, hi) table."""
    f64 = np.asarray(x).dtype == np.float64
    x = _f64(x) if f64 else _f32(x)
    assert x.size == lo * mid * hi
    y = np.empty(x.size, x.dtype)
    fn = lib().nt2ref_scan_lmh_f64 if f64 else lib().nt2ref_scan_lmh_f32
    fn(SCAN_OPS[op], _p(x), lo, mid, hi, _p(y))
    return y


def fold_dot_f32(x, y) -> np.float32:
    x, y = _f32(x), _f32(y)
    return np.float32(lib().nt2ref_fold_dot_f32(_p(x), _p(y), x.size))


def exact_sum_f32(x) -> tuple[float, float]:
    x = _f32(x)
    out = np.zeros(2)
    lib().nt2ref_exact_sum_f32(_p(x), x.size, _p(out))
    return float(out[0]), float(out[1])


def exact_dot_f32(x, y) -> tuple[float, float]:
    x, y = _f32(x), _f32(y)
    out = np.zeros(2)
    lib().nt2ref_exact_dot_f32(_p(x), _p(y), x.size, _p(out))
    return float(out[0]), float(out[1])


def bs_fwd(S, K, T, r, sig, dtype=np.float32):
    conv = _f32 if dtype == np.float32 else _f64
    S, K, T, r, sig = (conv(a) for a in (S, K, T, r, sig))
    call, put = np.empty_like(S), np.empty_like(S)
    fn = lib().nt2ref_bs_fwd_f32 if dtype == np.float32 else lib().nt2ref_bs_fwd_f64
    fn(_p(S), _p(K), _p(T), _p(r), _p(sig), _p(call), _p(put), S.size)
    return call, put


def bs_fwd_adj(S, K, T, r, sig, w, dtype=np.float32):
    """Returns (price, dS, dK, dT, dr, dsig, totals[2]) in dtype."""
    conv = _f32 if dtype == np.float32 else _f64
    ins = [conv(a) for a in (S, K, T, r, sig, w)]
    outs = [np.empty_like(ins[0]) for _ in range(6)]
    tot = np.zeros(2)
    fn = lib().nt2ref_bs_fwd_adj_f32 if dtype == np.float32 else lib().nt2ref_bs_fwd_adj_f64
    fn(*[_p(a) for a in ins], *[_p(a) for a in outs], _p(tot), ins[0].size)
    return (*outs, tot)


def cpu_bs_fwd_adj_f32(ins, outs, l2_bytes=1 << 20):
    """Multithreaded CPU baseline (NT2 shared-memory grain policy); arrays preallocated."""
    lib().nt2ref_cpu_bs_fwd_adj_f32(*[_p(a) for a in ins], *[_p(a) for a in outs], ins[0].size, l2_bytes)


def cpu_simd_bs_fwd_adj_f32(ins, outs, l2_bytes=1 << 20):
    """The vectorised (branch-free, AVX2) CPU baseline under the same grain policy."""
    lib().nt2ref_cpu_simd_bs_fwd_adj_f32(*[_p(a) for a in ins], *[_p(a) for a in outs], ins[0].size, l2_bytes)


def cpu_simd_isa() -> str:
    """Instruction set of the vectorised CPU clones on this host."""
    return lib().nt2ref_cpu_simd_isa().decode()


def cpu_simd_map_f32(op: str, x) -> np.ndarray:
    """Branch-free float exp, log, erfc, normcdf, normpdf of the vectorised baseline."""
    x = _f32(x)
    y = np.empty_like(x)
    lib().nt2ref_cpu_simd_map_f32(OPS[op], _p(x), _p(y), x.size)
    return y


def colfold_f32(X, s, k, p) -> np.ndarray:
    """X column-major (rows, cols) as a Fortran-ordered float32 array."""
    X = np.asfortranarray(X, dtype=np.float32)
    rows, cols = X.shape
    s, k, p = _f32(s), _f32(k), _f32(p)
    z = np.empty(cols, np.float32)
    lib().nt2ref_colfold_f32(_p(X), rows, cols, rows, _p(s), _p(k), _p(p), _p(z))
    return z


def colfold_gc_f64(X, s, k, p, g=None):
    X = np.asfortranarray(X, dtype=np.float32)
    rows, cols = X.shape
    s, k, p = _f32(s), _f32(k), _f32(p)
    gp = None
    if g is not None:
        g = _f32(g)
        gp = _p(g)
    outs = [np.empty(cols, np.float64) for _ in range(4)]
    lib().nt2ref_colfold_gc_f64(_p(X), rows, cols, rows, _p(s), _p(k), _p(p), gp, *[_p(o) for o in outs])
    return tuple(outs)


def ulp_stats_f32(got, ref) -> dict:
    """float results vs FP64 reference: ULP at ref (|ref| floored at FLT_MIN) and relative error."""
    got, ref = _f32(got), _f64(ref)
    out = np.zeros(4)
    lib().nt2ref_ulp_stats_f32(_p(got), _p(ref), got.size, _p(out))
    return {"max_ulp": out[0], "argmax": int(out[1]), "max_rel": out[2], "nonfinite_mismatch": int(out[3])}


def ulp_stats_f64(got, ref) -> dict:
    got, ref = _f64(got), _f64(ref)
    out = np.zeros(4)
    lib().nt2ref_ulp_stats_f64(_p(got), _p(ref), got.size, _p(out))
    return {"max_ulp": out[0], "argmax": int(out[1]), "max_rel": out[2], "nonfinite_mismatch": int(out[3])}


def cmp_f32(a, b) -> dict:
    """Two float arrays: max std-ULP (nextafter steps) and max NT2-ULP (half scale)."""
    a, b = _f32(a), _f32(b)
    out = np.zeros(4)
    lib().nt2ref_cmp_f32(_p(a), _p(b), a.size, _p(out))
    return {"max_std_ulp": out[0], "argmax": int(out[1]), "max_nt2_ulp": out[2], "nan_mismatch": int(out[3])}


def nt2_ulpdist(a, b) -> float:
    if isinstance(a, (np.float32,)) or isinstance(b, (np.float32,)):
        return lib().nt2ref_nt2_ulpdist_f32(float(a), float(b))
    return lib().nt2ref_nt2_ulpdist_f64(float(a), float(b))


def exhaustive_hist_f32(op: str, lo: float, hi: float) -> np.ndarray:
    h = np.zeros(130, np.uint64)
    lib().nt2ref_exhaustive_hist_f32(OPS[op], lo, hi, _p(h))
    return h
