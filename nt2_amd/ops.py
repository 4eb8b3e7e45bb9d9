"""Torch-facing wrappers over the libnt2hip C ABI.

Tensors must be contiguous CUDA (HIP) tensors of the stated dtype; each call
launches on torch's current stream.  These mirror NT2's table operations on
the path: ``Y = a*X + Y`` (saxpy), ``Y = f(X)`` (one functor node),
``globalsum``/``dot`` (whole-table inner_fold), the Black-Scholes loop nest
and its adjoint, and the column fold with its adjoint.
"""
from __future__ import annotations

import ctypes as C

from . import (DEFAULT, EXACT, F64_OPS, FAST, FOLD_OPS, FUSED, OPS, SCAN_OPS, SUM_KAHAN, call)
from .layout import streams


def _torch():
    import torch

    return torch


def _stream():
    return C.c_void_p(_torch().cuda.current_stream().cuda_stream)


def _ptr(t) -> C.c_void_p:
    return C.c_void_p(t.data_ptr()) if t is not None else C.c_void_p(0)


def _chk(t, dtype, name="tensor"):
    torch = _torch()
    if not t.is_cuda:
        raise ValueError(f"{name} must be a CUDA tensor")
    if t.dtype != dtype:
        raise TypeError(f"{name} must be {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    return t


def _same_size(what, *ts):
    n = ts[0].numel()
    for t in ts[1:]:
        if t.numel() != n:
            raise ValueError(f"{what}: tables differ in size ({n} vs {t.numel()})")


def saxpy_(a: float, x, y):
    """y <- a*x + y in place (two roundings, bit-identical to NT2's scalar path)."""
    torch = _torch()
    _chk(x, torch.float32, "x"), _chk(y, torch.float32, "y")
    if x.numel() != y.numel():
        raise ValueError(f"saxpy_: sizes differ ({x.numel()} vs {y.numel()})")
    call("nt2h_saxpy_f32", C.c_float(a), _ptr(x), _ptr(y), y.numel(), _stream())
    return y


def map(op: str, x, precision: int = EXACT, out=None):  # noqa: A001 (NT2 vocabulary)
    torch = _torch()
    if x.dtype == torch.float32:
        fn, code = "nt2h_map_f32", OPS[op]
    elif x.dtype == torch.float64:
        fn, code = "nt2h_map_f64", F64_OPS[op]
    else:
        raise TypeError("map: float32 or float64")
    _chk(x, x.dtype, "x")
    y = torch.empty_like(x) if out is None else _chk(out, x.dtype, "out")
    _same_size("map", x, y)
    call(fn, code, precision, _ptr(x), _ptr(y), x.numel(), _stream())
    return y


def bs_fwd(S, K, T, r, sig, precision: int = EXACT):
    """Black-Scholes call and put (5 input streams -> 2 output streams, one kernel)."""
    torch = _torch()
    dt = S.dtype
    for n, t in zip("SKTrs", (S, K, T, r, sig)):
        _chk(t, dt, n)
    _same_size("bs_fwd", S, K, T, r, sig)
    call_, put = streams(2, S.numel(), dt, S.device)
    fn = "nt2h_bs_fwd_f32" if dt == torch.float32 else "nt2h_bs_fwd_f64"
    call(fn, precision, *[_ptr(t) for t in (S, K, T, r, sig, call_, put)], S.numel(), _stream())
    return call_, put


def _fold(fn, tensors, n, extra, out=None):
    torch = _torch()
    res = torch.empty(1, dtype=torch.float64, device=tensors[0].device) if out is None else out
    call(fn, *[_ptr(t) for t in tensors], n, *extra, _ptr(res), _stream())
    return res


def sum(x, mode: int = SUM_KAHAN, out=None):  # noqa: A001 (NT2 vocabulary: sum / globalsum)
    """Whole-table sum; returns a 1-element float64 device tensor."""
    torch = _torch()
    if x.dtype == torch.float32:
        return _fold("nt2h_sum_f32", [_chk(x, torch.float32, "x")], x.numel(), [mode], out)
    _chk(x, torch.float64, "x")
    return _fold("nt2h_sum_f64", [x], x.numel(), [], out)


def dot(x, y, mode: int = SUM_KAHAN, out=None):
    torch = _torch()
    if x.numel() != y.numel():
        raise ValueError(f"dot: sizes differ ({x.numel()} vs {y.numel()})")
    if x.dtype == torch.float32:
        return _fold("nt2h_dot_f32", [_chk(x, torch.float32, "x"), _chk(y, torch.float32, "y")], x.numel(),
                     [mode], out)
    return _fold("nt2h_dot_f64", [_chk(x, torch.float64, "x"), _chk(y, torch.float64, "y")], x.numel(), [], out)


def bs_fwd_adj(S, K, T, r, sig, w, totals: bool = True, outs=None, tot_out=None):
    """price and the w-weighted input cotangents (dS, dK, dT, dr, dsig), plus
    totals = [sum dr, sum dsig] as a float64 device tensor (or None); the
    kernel writes every output, so ``outs``/``tot_out`` may be uninitialised."""
    torch = _torch()
    dt = S.dtype
    for n, t in zip(("S", "K", "T", "r", "sig", "w"), (S, K, T, r, sig, w)):
        _chk(t, dt, n)
    _same_size("bs_fwd_adj", S, K, T, r, sig, w)
    if outs is None:
        outs = streams(6, S.numel(), dt, S.device)
    elif len(outs) != 6:
        raise ValueError("bs_fwd_adj: outs must hold 6 tables")
    for n, t in zip(("price", "dS", "dK", "dT", "dr", "dsig"), outs):
        _chk(t, dt, n)
    _same_size("bs_fwd_adj", S, *outs)
    tot = None
    if totals:
        tot = torch.empty(2, dtype=torch.float64, device=S.device) if tot_out is None else tot_out
        _chk(tot, torch.float64, "tot_out")
        if tot.numel() != 2:
            raise ValueError("bs_fwd_adj: tot_out must hold 2 values")
    fn = "nt2h_bs_fwd_adj_f32" if dt == torch.float32 else "nt2h_bs_fwd_adj_f64"
    call(fn, *[_ptr(t) for t in (S, K, T, r, sig, w)], *[_ptr(o) for o in outs], _ptr(tot), S.numel(), _stream())
    return (*outs, tot)


def colfold_gc(X, s, k, p, gz=None, grad: bool = True, precision: int = FAST):
    """X: (cols, rows) row-major CUDA tensor == column-major (rows, cols) table.

    Returns z and, when grad, (dz_ds, dz_dk, dz_dp) for cotangent gz (None = ones).
    precision: FAST (hardware exp2, argument prescaled by 1/pi) or FUSED (NT2's
    exp and sin/cos algorithms with FMA)."""
    torch = _torch()
    _chk(X, torch.float32, "X")
    cols, rows = X.shape
    for n, t in zip("skp", (s, k, p)):
        _chk(t, torch.float32, n)
        if t.numel() != cols:
            raise ValueError(f"colfold_gc: {n} must hold one value per column ({cols})")
    if gz is not None:
        _chk(gz, torch.float32, "gz")
        if gz.numel() != cols:
            raise ValueError(f"colfold_gc: gz must hold one value per column ({cols})")
    z = torch.empty(cols, dtype=torch.float32, device=X.device)
    g = [torch.empty_like(z) for _ in range(3)] if grad else [None, None, None]
    call("nt2h_colfold_gc_f32_prec", precision, _ptr(X), rows, cols, rows, _ptr(s), _ptr(k), _ptr(p), _ptr(gz),
         _ptr(z), *[_ptr(t) for t in g], _stream())
    return (z, *g) if grad else z


def nt2_extents(t) -> list[int]:
    """NT2 extents (column-major) of a row-major torch tensor: reversed shape."""
    return list(reversed(t.shape)) or [1]


def lmh(extents, dim: int) -> tuple[int, int, int]:
    """The (lo, mid, hi) reshaping NT2 applies for a reduction or scan along
    1-based dimension ``dim`` of a table with the given extents."""
    if dim < 1:
        raise ValueError("dimensions are 1-based")
    ext = list(extents) + [1] * max(0, dim - len(extents))
    lo = 1
    for e in ext[: dim - 1]:
        lo *= e
    hi = 1
    for e in ext[dim:]:
        hi *= e
    return lo, ext[dim - 1], hi


def first_non_singleton(extents) -> int:
    for k, e in enumerate(extents, 1):
        if e != 1:
            return k
    return 1


def _fdt(x, what):
    """float32 / float64 table -> (dtype, ABI suffix)."""
    torch = _torch()
    if x.dtype == torch.float32:
        return torch.float32, "f32"
    if x.dtype == torch.float64:
        return torch.float64, "f64"
    raise TypeError(f"{what}: float32 or float64 table, got {x.dtype}")


def fold(op: str, x, dim: int | None = None, out=None):
    """NT2 reduction functor ``op`` (sum, prod, maximum, minimum, asum1, asum2,
    norm2, mean) of table ``x`` along 1-based NT2 dimension ``dim`` (default:
    the first non-singleton one).  ``x`` is a contiguous float32/float64 CUDA
    tensor whose reversed shape is the NT2 extent; the result has extent 1
    along dim."""
    torch = _torch()
    dt, sfx = _fdt(x, "fold")
    _chk(x, dt, "x")
    if op not in FOLD_OPS:
        raise ValueError(f"fold: unknown functor {op!r}")
    ext = nt2_extents(x)
    if dim is None:
        dim = first_non_singleton(ext)
    lo, mid, hi = lmh(ext, dim)
    oext = list(ext) + [1] * max(0, dim - len(ext))
    oext[dim - 1] = 1
    shape = tuple(reversed(oext))
    y = torch.empty(shape, dtype=dt, device=x.device) if out is None else _chk(out, dt, "out")
    if y.numel() != lo * hi:
        raise ValueError(f"fold: out must hold {lo * hi} values")
    call("nt2h_fold_" + sfx, FOLD_OPS[op], _ptr(x), lo, mid, hi, _ptr(y), _stream())
    return y


def global_fold(op: str, x, out=None):
    """globalX(x): the functor over every element (a 1-element tensor of x's type)."""
    torch = _torch()
    dt, sfx = _fdt(x, "global_fold")
    _chk(x, dt, "x")
    y = torch.empty(1, dtype=dt, device=x.device) if out is None else _chk(out, dt, "out")
    call("nt2h_fold_" + sfx, FOLD_OPS[op], _ptr(x), 1, x.numel(), 1, _ptr(y), _stream())
    return y


def scan(op: str, x, dim: int | None = None, out=None):
    """NT2 scan ``op`` (cumsum, cumprod, cummax, cummin) of table ``x`` along
    1-based NT2 dimension ``dim`` (default: first non-singleton)."""
    torch = _torch()
    dt, sfx = _fdt(x, "scan")
    _chk(x, dt, "x")
    if op not in SCAN_OPS:
        raise ValueError(f"scan: unknown functor {op!r}")
    ext = nt2_extents(x)
    if dim is None:
        dim = first_non_singleton(ext)
    lo, mid, hi = lmh(ext, dim)
    y = torch.empty_like(x) if out is None else _chk(out, dt, "out")
    _same_size("scan", x, y)
    call("nt2h_scan_" + sfx, SCAN_OPS[op], _ptr(x), lo, mid, hi, _ptr(y), _stream())
    return y


__all__ = ["saxpy_", "map", "bs_fwd", "sum", "dot", "bs_fwd_adj", "colfold_gc", "fold", "global_fold", "scan",
           "lmh", "EXACT", "FUSED", "DEFAULT"]
