"""The C ABI library: loads, exports exactly what include/nt2hip.h declares,
and rejects bad arguments before touching the device (no GPU needed)."""
import ctypes as C
import os
import re

import pytest

import nt2_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "nt2hip.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nt2h_[a-z0-9_]+)\s*\(", text)))


def test_library_loads_and_exports_every_declared_symbol():
    L = nt2_amd.lib()
    names = declared()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # and the Python binding knows the signature of each
    assert set(names) == set(nt2_amd.SIGNATURES)


def test_strerror_and_no_device():
    L = nt2_amd.lib()
    assert L.nt2h_strerror(0) == b"success"
    assert L.nt2h_strerror(-2) == b"null pointer"
    assert L.nt2h_strerror(12345) == b"unknown status"


P = C.c_void_p


def fake(n=1):
    """A non-null address that is never dereferenced (validation fails first)."""
    return P(0x1000)


def test_elementwise_argument_checks():
    L = nt2_amd.lib()
    assert L.nt2h_saxpy_f32(C.c_float(2.5), None, fake(), 8, None) == -2
    assert L.nt2h_saxpy_f32(C.c_float(2.5), fake(), fake(), -1, None) == -1
    assert L.nt2h_saxpy_f32(C.c_float(2.5), None, None, 0, None) == 0  # empty: nothing to do
    assert L.nt2h_map_f32(99, 0, fake(), fake(), 8, None) == -3
    assert L.nt2h_map_f64(10, 0, fake(), fake(), 8, None) == -3  # fastnormcdf is float only
    assert L.nt2h_map_f32(0, 7, fake(), fake(), 8, None) == -1  # bad precision
    assert L.nt2h_map_f32(0, 0, None, fake(), 8, None) == -2
    assert L.nt2h_map_f32(0, 0, None, None, 0, None) == 0
    args = [fake()] * 7
    assert L.nt2h_bs_fwd_f32(0, *args[:6], None, 8, None) == -2
    assert L.nt2h_bs_fwd_f32(3, *args, 8, None) == -1
    assert L.nt2h_bs_fwd_f64(0, *args, -5, None) == -1


def test_fold_and_adjoint_argument_checks():
    L = nt2_amd.lib()
    assert L.nt2h_sum_f32(fake(), 8, 9, fake(), None) == -1  # bad mode
    assert L.nt2h_sum_f32(fake(), 8, 1, None, None) == -2  # no output
    assert L.nt2h_sum_f32(None, 8, 1, fake(), None) == -2
    assert L.nt2h_dot_f32(fake(), None, 8, 1, fake(), None) == -2
    assert L.nt2h_sum_f64(fake(), -1, fake(), None) == -1
    a12 = [fake()] * 12
    assert L.nt2h_bs_fwd_adj_f32(*a12[:11], None, None, 8, None) == -2
    assert L.nt2h_bs_fwd_adj_f32(*a12, None, -8, None) == -1
    assert L.nt2h_bs_fwd_adj_f64(None, *a12[1:], None, 8, None) == -2


def test_colfold_argument_checks():
    L = nt2_amd.lib()
    f = fake()
    # ld < rows
    assert L.nt2h_colfold_gc_f32(f, 10, 4, 8, f, f, f, None, f, f, f, f, None) == -1
    assert L.nt2h_colfold_gc_f32(f, -1, 4, 8, f, f, f, None, f, f, f, f, None) == -1
    # some but not all gradient pointers
    assert L.nt2h_colfold_gc_f32(f, 10, 4, 10, f, f, f, None, f, f, None, f, None) == -2
    assert L.nt2h_colfold_gc_f32(f, 10, 4, 10, None, f, f, None, f, None, None, None, None) == -2
    assert L.nt2h_colfold_gc_f32(None, 10, 4, 10, f, f, f, None, f, None, None, None, None) == -2
    assert L.nt2h_colfold_gc_f32(f, 10, 0, 10, f, f, f, None, f, None, None, None, None) == 0  # no columns


def test_fold_scan_argument_checks():
    L = nt2_amd.lib()
    f = fake()
    assert L.nt2h_fold_f32(8, f, 1, 4, 1, f, None) == -3  # unknown functor
    assert L.nt2h_fold_f32(-1, f, 1, 4, 1, f, None) == -3
    assert L.nt2h_fold_f32(0, f, -1, 4, 1, f, None) == -1
    assert L.nt2h_fold_f32(0, None, 1, 4, 1, f, None) == -2
    assert L.nt2h_fold_f32(0, f, 1, 4, 1, None, None) == -2
    assert L.nt2h_fold_f32(0, None, 0, 4, 1, None, None) == 0  # no outputs
    assert L.nt2h_fold_f32(0, f, 1 << 40, 1 << 40, 1, f, None) == -7  # size overflow
    assert L.nt2h_scan_f32(4, f, 1, 4, 1, f, None) == -3
    assert L.nt2h_scan_f32(0, f, 1, -4, 1, f, None) == -1
    assert L.nt2h_scan_f32(0, f, 1, 4, 1, None, None) == -2
    assert L.nt2h_scan_f32(0, None, 1, 0, 1, None, None) == 0  # empty


def test_transfer_argument_checks():
    L = nt2_amd.lib()
    f = fake()
    assert L.nt2h_upload(None, f, 16, None) == -2
    assert L.nt2h_download(f, None, 16, None) == -2
    assert L.nt2h_upload(None, None, 0, None) == 0
    a12 = [f] * 12
    assert L.nt2h_bs_fwd_adj_f32_host(*a12[:11], None, None, 8, 0) == -2
    assert L.nt2h_bs_fwd_adj_f32_host(*a12, None, -8, 0) == -1


def test_runtime_argument_checks():
    L = nt2_amd.lib()
    assert L.nt2h_malloc(None, 16) == -2
    assert L.nt2h_memcpy(fake(), fake(), 16, 42, None) == -1
    assert L.nt2h_memcpy(None, fake(), 16, 1, None) == -2
    assert L.nt2h_memcpy(None, None, 0, 1, None) == 0
    assert L.nt2h_free(None) == 0
    cnt = C.c_int(-1)
    assert L.nt2h_device_count(C.byref(cnt)) == 0


def test_python_binding_raises_with_status():
    with pytest.raises(nt2_amd.NT2HError) as e:
        nt2_amd.call("nt2h_map_f32", 99, 0, None, None, 4, None)
    assert e.value.code == -3
