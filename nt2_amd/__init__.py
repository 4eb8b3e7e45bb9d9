"""nt2_amd — MI355X (gfx950) evaluator for NT2 table expressions.

The product is ``nt2_amd/lib/libnt2hip.so`` with the C ABI declared in
``include/nt2hip.h``; C++ callers use that header (or the Matlab-like front
end in ``include/nt2hip/table.hpp``).  This module is the Python binding of
the same ABI over ctypes, working on torch CUDA tensors (torch supplies
device memory and the current stream; every computation runs in the HIP
library).  There is no CPU fallback: if the library is missing or no GPU is
visible, calls raise.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# NT2H_LIB selects another build of the same sources (the sanitizer build of
# the host code)
LIB_PATH = os.environ.get("NT2H_LIB") or os.path.join(HERE, "lib", "libnt2hip.so")
CSRC = os.path.join(HERE, "csrc")

# enums of include/nt2hip.h
OPS = {
    "exp": 0, "log": 1, "sqrt": 2, "erf": 3, "erfc": 4, "normcdf": 5, "tanh": 6,
    "sin": 7, "cos": 8, "normpdf": 9, "fastnormcdf": 10, "exp2": 11, "exp10": 12,
}
F64_OPS = {k: v for k, v in OPS.items() if v <= 9}
EXACT, FUSED, FAST = 0, 1, 2
SUM_NAIVE, SUM_KAHAN, SUM_PAIRWISE = 0, 1, 2
H2H, H2D, D2H, D2D, DEFAULT = 0, 1, 2, 3, 4
FOLD_OPS = {"sum": 0, "prod": 1, "maximum": 2, "minimum": 3, "asum1": 4, "asum2": 5, "norm2": 6, "mean": 7}
SCAN_OPS = {"cumsum": 0, "cumprod": 1, "cummax": 2, "cummin": 3}

STATUS = {
    0: "NT2H_OK", -1: "NT2H_ERR_INVALID_ARG", -2: "NT2H_ERR_NULL_POINTER", -3: "NT2H_ERR_BAD_OP",
    -4: "NT2H_ERR_NO_DEVICE", -5: "NT2H_ERR_HIP", -6: "NT2H_ERR_OUT_OF_MEMORY", -7: "NT2H_ERR_TOO_LARGE",
    -8: "NT2H_ERR_CAPTURE", -9: "NT2H_ERR_WAIT",
}

# every exported entry point with its ctypes signature (restype, argtypes)
_P, _I, _I64, _F, _D, _SZ = C.c_void_p, C.c_int, C.c_int64, C.c_float, C.c_double, C.c_size_t
SIGNATURES = {
    "nt2h_init": (_I, [_I]),
    "nt2h_device_count": (_I, [_P]),
    "nt2h_strerror": (C.c_char_p, [_I]),
    "nt2h_last_hip_error": (_I, []),
    "nt2h_malloc": (_I, [_P, _SZ]),
    "nt2h_free": (_I, [_P]),
    "nt2h_host_alloc": (_I, [_P, _SZ]),
    "nt2h_host_free": (_I, [_P]),
    "nt2h_memcpy": (_I, [_P, _P, _SZ, _I, _P]),
    "nt2h_sync": (_I, [_P]),
    "nt2h_release_stream": (_I, [_P]),
    "nt2h_state_bytes": (_I, [_P]),
    "nt2h_stream_buffer": (_I, [_P, _I, _SZ, _P]),
    "nt2h_test_set_lookback_wait": (_I, [C.c_uint, _I]),
    "nt2h_saxpy_f32": (_I, [_F, _P, _P, _I64, _P]),
    "nt2h_map_f32": (_I, [_I, _I, _P, _P, _I64, _P]),
    "nt2h_map_f64": (_I, [_I, _I, _P, _P, _I64, _P]),
    "nt2h_bs_fwd_f32": (_I, [_I] + [_P] * 7 + [_I64, _P]),
    "nt2h_bs_fwd_f64": (_I, [_I] + [_P] * 7 + [_I64, _P]),
    "nt2h_sum_f32": (_I, [_P, _I64, _I, _P, _P]),
    "nt2h_dot_f32": (_I, [_P, _P, _I64, _I, _P, _P]),
    "nt2h_sum_f64": (_I, [_P, _I64, _P, _P]),
    "nt2h_dot_f64": (_I, [_P, _P, _I64, _P, _P]),
    "nt2h_bs_fwd_adj_f32": (_I, [_P] * 13 + [_I64, _P]),
    "nt2h_bs_fwd_adj_f64": (_I, [_P] * 13 + [_I64, _P]),
    "nt2h_bs_fwd_adj_f32_noalloc": (_I, [_P] * 13 + [_I64, _P]),
    "nt2h_colfold_gc_f32": (_I, [_P, _I64, _I64, _I64] + [_P] * 8 + [_P]),
    "nt2h_colfold_gc_f32_prec": (_I, [_I, _P, _I64, _I64, _I64] + [_P] * 8 + [_P]),
    "nt2h_fold_f32": (_I, [_I, _P, _I64, _I64, _I64, _P, _P]),
    "nt2h_fold_f64": (_I, [_I, _P, _I64, _I64, _I64, _P, _P]),
    "nt2h_scan_f64": (_I, [_I, _P, _I64, _I64, _I64, _P, _P]),
    "nt2h_upload": (_I, [_P, _P, _SZ, _P]),
    "nt2h_download": (_I, [_P, _P, _SZ, _P]),
    "nt2h_bs_fwd_adj_f32_host": (_I, [_P] * 13 + [_I64, _I64]),
    "nt2h_scan_f32": (_I, [_I, _P, _I64, _I64, _I64, _P, _P]),
}


class NT2HError(RuntimeError):
    def __init__(self, fn: str, code: int):
        super().__init__(f"{fn} failed: {STATUS.get(code, code)} ({code})")
        self.code = code


_lib = None


def build(force: bool = False) -> str:
    """Compile libnt2hip.so for gfx950 with hipcc (csrc/Makefile)."""
    import subprocess

    if force:
        subprocess.run(["make", "-s", "-C", CSRC, "clean"], check=True)
    jobs = str(min(8, os.cpu_count() or 1))
    subprocess.run(["make", "-s", "-j", jobs, "-C", CSRC], check=True)
    return LIB_PATH


def lib() -> C.CDLL:
    """Load libnt2hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with nt2_amd.build() (hipcc, gfx950)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(fn: str, code: int) -> None:
    if code != 0:
        raise NT2HError(fn, code)


def call(fn: str, *args) -> None:
    check(fn, getattr(lib(), fn)(*args))


from . import ops  # noqa: E402,F401  (torch-facing wrappers)
