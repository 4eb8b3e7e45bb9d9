"""GPU parity of the other reduction functors and the scans (nt2h_fold_f32,
nt2h_scan_f32) against the oracle's sequential NT2 restatement.

Bars: maximum/minimum/cummax/cummin bit-identical to NT2 (NaN, signed zero
and neutrals included); the compensated sum family within 1 float ULP of the
exactly rounded FP64 result and inside NT2's own sequential error; prod and
cumprod within 2 ULP of the FP64 product; the NT2 unit vectors exactly.
"""
import numpy as np
import pytest

import nt2_amd
from nt2_amd import ops

pytestmark = pytest.mark.gpu

# (lo, mid, hi): run kernel single segment, run kernel segmented, short runs
# (column kernel), outer folds with and without segments, NT2's 3x5x2 table,
# and lo % 4 == 0 (the 16-byte column kernels) with and without segments
SHAPES = [(1, 5000, 37), (1, (1 << 20) + 3, 1), (1, 1 << 18, 2), (1, 100, 1000), (1, 700, 3), (7, 3000, 5),
          (1000, 3, 7), (3, 5, 2), (2, 40000, 1), (1, 1, 9), (16, 5000, 3), (1024, 300, 2),
          # narrow tables: 1, 4, 16 and 15 groups of four outputs (row-phase tile kernels)
          (4, 30000, 1), (8, 20000, 2), (64, 4097, 1), (12, 777, 5), (12, 5000, 1)]


def _dev_tensor(a, dev):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _sfx(x):
    return "f64" if np.asarray(x).dtype == np.float64 else "f32"


def _fold(op, x, lo, mid, hi, dev):
    import torch

    gx = _dev_tensor(x, dev)
    out = torch.empty(lo * hi, dtype=gx.dtype, device=dev)
    nt2_amd.call("nt2h_fold_" + _sfx(x), nt2_amd.FOLD_OPS[op], gx.data_ptr(), lo, mid, hi, out.data_ptr(),
                 torch.cuda.current_stream().cuda_stream)
    return out.cpu().numpy()


def _scan(op, x, lo, mid, hi, dev):
    import torch

    gx = _dev_tensor(x, dev)
    y = torch.empty_like(gx)
    nt2_amd.call("nt2h_scan_" + _sfx(x), nt2_amd.SCAN_OPS[op], gx.data_ptr(), lo, mid, hi, y.data_ptr(),
                 torch.cuda.current_stream().cuda_stream)
    return y.cpu().numpy()


def _specials_data(n, seed):
    """Random values with NaN, +-0, +-inf and +-FLT_MAX sprinkled in."""
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(n).astype(np.float32)
    sp = np.array([np.nan, 0.0, -0.0, np.inf, -np.inf, 3.4028235e38, -3.4028235e38], np.float32)
    k = rng.random(n) < 0.02
    x[k] = sp[rng.integers(0, sp.size, int(k.sum()))]
    return x


def _bits_eq(a, b):
    u = np.uint64 if a.dtype == np.float64 else np.uint32
    return (a.view(u) == b.view(u)) | (np.isnan(a) & np.isnan(b))


@pytest.mark.parametrize("op", ["maximum", "minimum"])
@pytest.mark.parametrize("shape", SHAPES)
def test_extreme_folds_bit_identical(dev, ref, op, shape):
    lo, mid, hi = shape
    for seed, data in ((1, "specials"), (2, "normal"), (3, "few-values")):
        if data == "specials":
            x = _specials_data(lo * mid * hi, seed)
        elif data == "normal":
            x = np.random.default_rng(seed).standard_normal(lo * mid * hi).astype(np.float32)
        else:  # many ties, signed zeros
            x = np.random.default_rng(seed).choice(np.array([0.0, -0.0, 1.0, -1.0], np.float32), lo * mid * hi)
        got = _fold(op, x, lo, mid, hi, dev)
        want = ref.fold_lmh(op, x, lo, mid, hi)
        eq = _bits_eq(got, want)
        assert eq.all(), (data, np.nonzero(~eq)[0][:5], got[~eq][:5], want[~eq][:5])


def test_extreme_fold_nan_before_the_tail(dev, ref):
    """The overall maximum before the last NaN: the tail is rescanned."""
    for n in (1 << 20, 5000, 100):
        x = np.random.default_rng(4).random(n).astype(np.float32)
        x[3] = 7.0  # the largest value, early
        x[n // 2] = np.nan  # the last NaN, after it
        for op in ("maximum", "minimum"):
            for lo, mid, hi in ((1, n, 1), (n, 1, 1)):
                got = _fold(op, x, lo, mid, hi, dev)
                want = ref.fold_lmh(op, x, lo, mid, hi)
                assert _bits_eq(got, want).all()
        x[-1] = np.nan
        assert np.isnan(_fold("maximum", x, 1, n, 1, dev)[0])
    # one workgroup per run (single segment): each run has its extreme before a NaN
    lo, mid, hi = 1, 600, 20
    x = np.random.default_rng(12).random(lo * mid * hi).astype(np.float32)
    for h in range(hi):
        x[h * mid + h] = 5.0
        x[h * mid + 300 + h] = np.nan
    x[3 * mid:4 * mid] = np.nan  # an all-NaN run
    for op in ("maximum", "minimum"):
        got = _fold(op, x, lo, mid, hi, dev)
        assert _bits_eq(got, ref.fold_lmh(op, x, lo, mid, hi)).all()


def _exact_sums(x, lo, mid, hi, kind):
    v = x.astype(np.float64).reshape(hi, mid, lo)  # C order of (h, m, i)
    if kind == "asum1":
        v = np.abs(v)
    elif kind in ("asum2", "norm2"):
        v = v * v
    return v.sum(axis=1).reshape(-1)  # index i + lo*h


@pytest.mark.parametrize("op", ["sum", "asum1", "asum2", "norm2", "mean"])
@pytest.mark.parametrize("shape", SHAPES)
def test_sum_family_accuracy(dev, ref, op, shape):
    lo, mid, hi = shape
    x = np.random.default_rng(5).standard_normal(lo * mid * hi).astype(np.float32) + np.float32(0.25)
    got = _fold(op, x, lo, mid, hi, dev)
    tot = _exact_sums(x, lo, mid, hi, op)
    t32 = tot.astype(np.float32)
    if op == "norm2":
        want = np.sqrt(t32)
    elif op == "mean":
        want = (np.float32(1) / np.float32(mid)) * t32  # rec(size) * sum, as NT2
    else:
        want = t32
    # within 1 ULP of NT2's last steps applied to the exactly rounded total
    ulp = np.spacing(np.abs(want).astype(np.float32))
    mag = _exact_sums(np.abs(x), lo, mid, hi, "asum1" if op != "asum2" and op != "norm2" else op)
    assert np.all(np.abs(got.astype(np.float64) - want) <= ulp + 1e-14 * mag), op
    # and NT2's sequential float fold is no closer than its own error bound
    seq = ref.fold_lmh(op, x, lo, mid, hi)
    gamma = mid * np.finfo(np.float32).eps
    scale = mag * (1.0 / mid if op == "mean" else 1.0)
    if op == "norm2":
        scale = np.sqrt(mag)
    assert np.all(np.abs(seq.astype(np.float64) - got) <= gamma * scale + 2 * ulp)


@pytest.mark.parametrize("shape", SHAPES)
def test_prod(dev, ref, shape):
    lo, mid, hi = shape
    rng = np.random.default_rng(6)
    x = rng.uniform(0.9, 1.1, lo * mid * hi).astype(np.float32)
    got = _fold("prod", x, lo, mid, hi, dev)
    want = np.prod(x.astype(np.float64).reshape(hi, mid, lo), axis=1).reshape(-1)
    assert np.all(np.abs(got - want) <= 2 * np.spacing(want.astype(np.float32)))
    # exact powers of two: identical to NT2's sequential product
    p = rng.choice(np.array([0.5, 2.0, -2.0, 1.0], np.float32), lo * mid * hi)
    if mid <= 100:
        assert np.array_equal(_fold("prod", p, lo, mid, hi, dev), ref.fold_lmh("prod", p, lo, mid, hi))


def test_prod_specials_and_range(dev, ref):
    x = np.array([2.0] * 200 + [0.5] * 200, np.float32)  # 2^200 mid-way: NT2's float product is inf
    assert _fold("prod", x, 1, 400, 1, dev)[0] == 1.0
    assert np.isinf(ref.fold_lmh("prod", x, 1, 400, 1)[0])  # the documented difference
    for v in ([0.0, np.inf, 1.0], [np.inf, 0.0], [np.nan, 1.0]):
        a = np.array(v, np.float32)
        assert np.isnan(_fold("prod", a, 1, a.size, 1, dev)[0])
    z = np.array([-1.0, 0.0, 3.0], np.float32)
    g = _fold("prod", z, 1, 3, 1, dev)[0]
    assert g == 0 and np.signbit(g) == np.signbit(ref.fold_lmh("prod", z, 1, 3, 1)[0])


def test_empty_middle_gives_neutrals(dev):
    import torch

    out = torch.full((4,), 7.0, device=dev)
    want = {"sum": 0.0, "prod": 1.0, "maximum": -3.4028235e38, "minimum": 3.4028235e38, "norm2": 0.0}
    for op, v in want.items():
        nt2_amd.call("nt2h_fold_f32", nt2_amd.FOLD_OPS[op], 0, 2, 0, 2, out.data_ptr(),
                     torch.cuda.current_stream().cuda_stream)
        assert np.all(out.cpu().numpy() == np.float32(v)), op
    nt2_amd.call("nt2h_fold_f32", nt2_amd.FOLD_OPS["mean"], 0, 2, 0, 2, out.data_ptr(),
                 torch.cuda.current_stream().cuda_stream)
    assert np.all(np.isnan(out.cpu().numpy()))


def test_fold_unit_tables(dev):
    """NT2 unit tables through the torch-facing wrapper (reversed shapes)."""
    import torch

    y = torch.tensor([[i + 10 * j for i in range(1, 6)] for j in range(1, 4)], dtype=torch.float32, device=dev)
    assert ops.fold("maximum", y, 1).cpu().numpy().ravel().tolist() == [15, 25, 35]
    assert ops.fold("maximum", y, 2).cpu().numpy().ravel().tolist() == [31, 32, 33, 34, 35]
    assert ops.global_fold("maximum", y).item() == 35
    a = torch.arange(1, 5, dtype=torch.float32, device=dev)
    assert ops.global_fold("prod", a).item() == 24


# ------------------------------------------------------------------- scans
@pytest.mark.parametrize("op", ["cummax", "cummin"])
@pytest.mark.parametrize("shape", SHAPES)
def test_extreme_scans_bit_identical(dev, ref, op, shape):
    lo, mid, hi = shape
    for seed in (7, 8):
        x = _specials_data(lo * mid * hi, seed)
        got = _scan(op, x, lo, mid, hi, dev)
        want = ref.scan_lmh(op, x, lo, mid, hi)
        eq = _bits_eq(got, want)
        assert eq.all(), (np.nonzero(~eq)[0][:5], got[~eq][:5], want[~eq][:5])


@pytest.mark.parametrize("shape", SHAPES)
def test_cumsum_accuracy(dev, ref, shape):
    lo, mid, hi = shape
    rng = np.random.default_rng(9)
    x = rng.random(lo * mid * hi).astype(np.float32)  # positive: relative bound
    got = _scan("cumsum", x, lo, mid, hi, dev)
    want = np.cumsum(x.astype(np.float64).reshape(hi, mid, lo), axis=1).reshape(-1)
    w32 = want.astype(np.float32)
    assert np.all(np.abs(got - want) <= np.spacing(w32)), np.max(np.abs(got - want) / np.spacing(w32))
    # mixed signs: absolute bound against the running sum of magnitudes
    z = rng.standard_normal(lo * mid * hi).astype(np.float32)
    gz = _scan("cumsum", z, lo, mid, hi, dev)
    wz = np.cumsum(z.astype(np.float64).reshape(hi, mid, lo), axis=1).reshape(-1)
    mag = np.cumsum(np.abs(z.astype(np.float64)).reshape(hi, mid, lo), axis=1).reshape(-1)
    assert np.all(np.abs(gz - wz) <= np.spacing(np.abs(wz).astype(np.float32)) + 1e-12 * mag)
    # NT2's sequential float recurrence stays inside its own error bound of ours
    seq = ref.scan_lmh("cumsum", z, lo, mid, hi)
    assert np.all(np.abs(seq - gz) <= mid * np.finfo(np.float32).eps * mag + np.spacing(np.abs(wz).astype(np.float32)))


@pytest.mark.parametrize("shape", SHAPES)
def test_cumprod_accuracy(dev, shape):
    lo, mid, hi = shape
    x = np.random.default_rng(10).uniform(0.99, 1.01, lo * mid * hi).astype(np.float32)
    got = _scan("cumprod", x, lo, mid, hi, dev)
    want = np.cumprod(x.astype(np.float64).reshape(hi, mid, lo), axis=1).reshape(-1)
    assert np.all(np.abs(got - want) <= 2 * np.spacing(want.astype(np.float32)))


@pytest.mark.parametrize("shape", [(1, (1 << 20) + 3, 1), (1, 5000, 37), (16, 5000, 3), (4, 30000, 1)])
def test_cumprod_powers_of_two_exact(dev, ref, shape):
    """Long products of 0.5 and of 2: every prefix is an exact power of two
    down to the smallest subnormal and then 0 (or up to FLT_MAX's exponent
    and then inf) — NT2's sequential float recurrence gives the same bits,
    however the products are grouped inside tiles and across segments."""
    lo, mid, hi = shape
    for v in (0.5, 2.0):
        x = np.full(lo * mid * hi, v, np.float32)
        got = _scan("cumprod", x, lo, mid, hi, dev)
        want = ref.scan_lmh("cumprod", x, lo, mid, hi)
        assert _bits_eq(got, want).all(), v


@pytest.mark.parametrize("shape", [(1, (1 << 26) + 12, 1), (1, (1 << 26) + 5, 1), (1, 1 << 26, 2)])
def test_long_run_scans(dev, ref, shape):
    """Runs of 2^26 elements and more (three passes over thousands of
    segments): 16-byte and scalar bodies, a partial last segment, two runs.
    cummax/cummin bit-identical to NT2's recurrence, cumsum/cumprod within
    the same bounds as the other shapes, and two calls give the same bits."""
    lo, mid, hi = shape
    n = lo * mid * hi
    x = _specials_data(n, 31)
    for op in ("cummax", "cummin"):
        got = _scan(op, x, lo, mid, hi, dev)
        assert _bits_eq(got, ref.scan_lmh(op, x, lo, mid, hi)).all(), op
    rng = np.random.default_rng(32)
    z = rng.standard_normal(n).astype(np.float32)
    gz = _scan("cumsum", z, lo, mid, hi, dev)
    # extended-precision reference: a float64 running sum over 2^26 terms
    # drifts by ~1e-12 of the magnitudes itself
    zl = z.astype(np.longdouble).reshape(hi, mid, lo)
    wz = np.cumsum(zl, axis=1).reshape(-1)
    mag = np.cumsum(np.abs(zl), axis=1).reshape(-1).astype(np.float64)
    del zl
    wz = wz.astype(np.float64)
    assert np.all(np.abs(gz - wz) <= np.spacing(np.abs(wz).astype(np.float32)) + 1e-12 * mag)
    assert _bits_eq(_scan("cumsum", z, lo, mid, hi, dev), gz).all()
    del wz, mag
    p = rng.uniform(1 - 1e-7, 1 + 1e-7, n).astype(np.float32)
    gp = _scan("cumprod", p, lo, mid, hi, dev)
    wp = np.cumprod(p.astype(np.float64).reshape(hi, mid, lo), axis=1).reshape(-1)
    assert np.all(np.abs(gp - wp) <= 2 * np.spacing(wp.astype(np.float32)))


@pytest.mark.parametrize("shape", [(1, 1 << 19, 1), (1, (1 << 16) + 4, 3), (1, 20000, 1), (1, 16384, 2),
                                   (1, 1 << 19, 4),
                                   # either side of the single-launch limits: 128 tiles of
                                   # 8192 floats per run (64 for cumprod) and 512 in all
                                   (1, 1 << 20, 1), (1, (1 << 20) + 1, 1), (1, (1 << 19) + 1, 1),
                                   (1, 1 << 20, 4), (1, 1 << 20, 5), (1, 1 << 19, 9)])
def test_lookback_scans(dev, ref, shape):
    """Runs of 2^14 to 2^20 floats are scanned in one launch, tiles finding
    their prefix from the published state of the tiles before them (a
    partial last tile, several runs); one tile more per run or in all takes
    the three passes instead.  cummax/cummin bit-identical to NT2's
    recurrence, cumsum/cumprod within the bounds of the other paths, and the
    bits do not depend on the timing: three calls agree."""
    lo, mid, hi = shape
    n = lo * mid * hi
    x = _specials_data(n, 41)
    for op in ("cummax", "cummin"):
        got = _scan(op, x, lo, mid, hi, dev)
        assert _bits_eq(got, ref.scan_lmh(op, x, lo, mid, hi)).all(), op
    rng = np.random.default_rng(42)
    z = rng.standard_normal(n).astype(np.float32)
    gz = _scan("cumsum", z, lo, mid, hi, dev)
    wz = np.cumsum(z.astype(np.float64).reshape(hi, mid, lo), axis=1).reshape(-1)
    mag = np.cumsum(np.abs(z.astype(np.float64)).reshape(hi, mid, lo), axis=1).reshape(-1)
    assert np.all(np.abs(gz - wz) <= np.spacing(np.abs(wz).astype(np.float32)) + 1e-12 * mag)
    p = rng.uniform(0.999, 1.001, n).astype(np.float32)
    gp = _scan("cumprod", p, lo, mid, hi, dev)
    wp = np.cumprod(p.astype(np.float64).reshape(hi, mid, lo), axis=1).reshape(-1)
    assert np.all(np.abs(gp - wp) <= 2 * np.spacing(wp.astype(np.float32)))
    for _ in range(2):
        assert _bits_eq(_scan("cumsum", z, lo, mid, hi, dev), gz).all()
        assert _bits_eq(_scan("cumprod", p, lo, mid, hi, dev), gp).all()


@pytest.mark.parametrize("shape", [(3, (1 << 21) + 5, 1), (7, 1 << 19, 2)])
def test_few_scalar_outputs_many_segments(dev, ref, shape):
    """Outer folds and scans of tables with a few outputs per row that are not
    a multiple of four (the scalar column kernels), split into tens of
    thousands of segments: maximum/minimum/cummax/cummin bit-identical to
    NT2's recurrence, sum and cumsum within the bounds of the other paths."""
    lo, mid, hi = shape
    x = _specials_data(lo * mid * hi, 51)
    for op in ("maximum", "minimum"):
        assert _bits_eq(_fold(op, x, lo, mid, hi, dev), ref.fold_lmh(op, x, lo, mid, hi)).all(), op
    for op in ("cummax", "cummin"):
        assert _bits_eq(_scan(op, x, lo, mid, hi, dev), ref.scan_lmh(op, x, lo, mid, hi)).all(), op
    z = np.random.default_rng(52).standard_normal(lo * mid * hi).astype(np.float32)
    z64 = z.astype(np.float64).reshape(hi, mid, lo)
    gs = _fold("sum", z, lo, mid, hi, dev).astype(np.float64)
    ws, ms = z64.sum(axis=1).reshape(-1), np.abs(z64).sum(axis=1).reshape(-1)
    assert np.all(np.abs(gs - ws) <= np.spacing(np.abs(ws).astype(np.float32)) + 1e-12 * ms)
    gz = _scan("cumsum", z, lo, mid, hi, dev)
    wz = np.cumsum(z64, axis=1).reshape(-1)
    mag = np.cumsum(np.abs(z64), axis=1).reshape(-1)
    assert np.all(np.abs(gz - wz) <= np.spacing(np.abs(wz).astype(np.float32)) + 1e-12 * mag)


def test_scan_unit_vectors(dev, ref):
    """modules/core/swar/unit/table/cumsum.cpp: reshape(_(1, 30), 3, 5, 2) along 1 and 2."""
    a = np.arange(1, 31, dtype=np.float32)
    assert np.array_equal(_scan("cumsum", a, 1, 3, 10, dev), ref.scan_lmh("cumsum", a, 1, 3, 10))
    assert np.array_equal(_scan("cumsum", a, 3, 5, 2, dev), ref.scan_lmh("cumsum", a, 3, 5, 2))
    assert _scan("cumsum", a, 3, 5, 2, dev)[-1] == 120
    p = np.full(6, 2.0, np.float32)
    assert np.array_equal(_scan("cumprod", p, 1, 6, 1, dev), 2.0 ** np.arange(1, 7))


# ------------------------------------------------------------ double tables
F64_SHAPES = [(1, 5000, 37), (1, (1 << 19) + 3, 1), (1, 1 << 17, 2), (1, 100, 1000), (6, 3000, 5), (1024, 300, 2),
              (3, 5, 2), (2, 20000, 1), (10, 3001, 3)]


@pytest.mark.parametrize("shape", F64_SHAPES)
def test_f64_extreme_folds_and_scans_bit_identical(dev, ref, shape):
    lo, mid, hi = shape
    x = _specials_data(lo * mid * hi, 21).astype(np.float64)
    x[::97] = np.float64(1.7976931348623157e308)  # DBL_MAX, beyond any float
    for op in ("maximum", "minimum"):
        got, want = _fold(op, x, lo, mid, hi, dev), ref.fold_lmh(op, x, lo, mid, hi)
        assert got.dtype == np.float64 and _bits_eq(got, want).all(), op
    for op in ("cummax", "cummin"):
        got, want = _scan(op, x, lo, mid, hi, dev), ref.scan_lmh(op, x, lo, mid, hi)
        assert _bits_eq(got, want).all(), op
    inf = np.full(lo * mid * hi, -np.inf)
    assert np.all(_fold("maximum", inf, lo, mid, hi, dev) == -np.finfo(np.float64).max)  # Maxinit = -DBL_MAX


@pytest.mark.parametrize("shape", F64_SHAPES)
def test_f64_sums_prod_scans(dev, ref, shape):
    lo, mid, hi = shape
    rng = np.random.default_rng(22)
    x = rng.standard_normal(lo * mid * hi) + 0.25
    exact = np.array([float(np.sum(np.array(r, dtype=np.longdouble))) for r in
                      x.reshape(hi, mid, lo).transpose(0, 2, 1).reshape(-1, mid)]) if mid * lo * hi < 3e6 else None
    got = _fold("sum", x, lo, mid, hi, dev)
    mag = np.abs(x).reshape(hi, mid, lo).sum(axis=1).reshape(-1)
    if exact is not None:  # compensated: within 1 ULP of the (extended-precision) exact total
        assert np.all(np.abs(got - exact) <= np.spacing(np.abs(exact)) + 1e-30 * mag)
    seq = ref.fold_lmh("sum", x, lo, mid, hi)  # NT2's sequential double sum, inside its error bound
    assert np.all(np.abs(seq - got) <= mid * np.finfo(np.float64).eps * mag)
    nm = _fold("norm2", x, lo, mid, hi, dev)
    assert np.allclose(nm, np.sqrt((x * x).reshape(hi, mid, lo).sum(axis=1).reshape(-1)), rtol=1e-15)
    p = rng.uniform(0.99, 1.01, lo * mid * hi)
    gp = _fold("prod", p, lo, mid, hi, dev)
    assert np.allclose(gp, np.prod(p.reshape(hi, mid, lo), axis=1).reshape(-1), rtol=mid * 2e-16 + 1e-15)
    cs = _scan("cumsum", x, lo, mid, hi, dev)
    wc = np.cumsum(x.reshape(hi, mid, lo), axis=1).reshape(-1)
    cm = np.cumsum(np.abs(x).reshape(hi, mid, lo), axis=1).reshape(-1)
    assert np.all(np.abs(cs - wc) <= mid * np.finfo(np.float64).eps * cm)  # within numpy's own error
    cp = _scan("cumprod", p, lo, mid, hi, dev)
    assert np.allclose(cp, np.cumprod(p.reshape(hi, mid, lo), axis=1).reshape(-1), rtol=mid * 2e-16 + 1e-15)


def test_f64_wrappers_and_unit_tables(dev):
    import torch

    y = torch.tensor([[i + 10 * j for i in range(1, 6)] for j in range(1, 4)], dtype=torch.float64, device=dev)
    assert ops.fold("maximum", y, 1).cpu().numpy().ravel().tolist() == [15, 25, 35]
    assert ops.global_fold("prod", torch.arange(1, 5, dtype=torch.float64, device=dev)).item() == 24
    a = torch.arange(1, 31, dtype=torch.float64, device=dev)
    assert ops.scan("cumsum", a.reshape(10, 3)).cpu().numpy().ravel()[:9].tolist() == [1, 3, 6, 4, 9, 15, 7, 15, 24]


def test_scan_in_place(dev, ref):
    import torch

    x = _specials_data(70000, 11)
    g = _dev_tensor(x, dev)
    nt2_amd.call("nt2h_scan_f32", nt2_amd.SCAN_OPS["cummax"], g.data_ptr(), 1, 70000, 1, g.data_ptr(),
                 torch.cuda.current_stream().cuda_stream)
    assert _bits_eq(g.cpu().numpy(), ref.scan_lmh("cummax", x, 1, 70000, 1)).all()


# ------------------------------------------------- non-finite data in the sums
# Every segment gets one of these; NT2's plain float recurrence gives +-inf
# or NaN for each of them whatever the order, and the compensated kernels
# must too (an error term that meets inf is NaN, which must not reach the
# value).  "overflow": every value FLT_MAX/4 (DBL_MAX/4), so partial sums
# overflow and the total's rounding to the element type is inf.
NONFINITE = ("finite", "+inf", "-inf", "nan", "+inf-inf", "overflow")
NONFINITE_SHAPES = [(1, 5000, 37), (7, 3000, 5), (16, 5000, 3), (1000, 3, 7), (4, 30000, 1), (12, 777, 5),
                    (1, 100, 1000), (1024, 300, 2)]


def _nonfinite_segments(lo, mid, hi, dtype, seed=31):
    """(hi, mid, lo) data whose segment (h, i) is case NONFINITE[(h*lo+i) % 6]."""
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.5, 1.5, (hi, mid, lo)).astype(dtype)
    big = np.finfo(dtype).max / 4
    for h in range(hi):
        for i in range(lo):
            case = NONFINITE[(h * lo + i) % len(NONFINITE)]
            k = rng.integers(0, mid, 2)
            if case == "+inf":
                x[h, k[0], i] = np.inf
            elif case == "-inf":
                x[h, k[0], i] = -np.inf
            elif case == "nan":
                x[h, k[0], i] = np.nan
            elif case == "+inf-inf" and mid > 1:
                x[h, k[0], i] = np.inf
                x[h, (k[0] + 1 + k[1] % (mid - 1)) % mid, i] = -np.inf
            elif case == "overflow":
                x[h, :, i] = big
    return x


def _ieee_fold(op, x3, dtype):
    """NT2's functor on (hi, mid, lo) data along mid, as the plain IEEE
    recurrence gives it for these segments (order does not matter for them)."""
    with np.errstate(all="ignore"):
        w = x3.astype(np.float64)
        if op == "asum1":
            w = np.abs(w)
        elif op in ("asum2", "norm2"):
            w = w * w
        s = w.sum(axis=1).astype(dtype)  # the total rounded to the element type: inf past its range
        if op == "norm2":
            s = np.sqrt(s)
        elif op == "mean":
            s = (dtype(1) / dtype(x3.shape[1])) * s
    return s.reshape(-1)


def _same_nonfinite(got, want, rtol):
    got, want = np.asarray(got, np.float64), np.asarray(want, np.float64)
    assert (np.isnan(got) == np.isnan(want)).all(), (np.flatnonzero(np.isnan(got) != np.isnan(want))[:5])
    inf = np.isinf(want)
    assert (got[inf] == want[inf]).all(), (got[inf][got[inf] != want[inf]][:5])
    fin = np.isfinite(want)
    assert np.isfinite(got[fin]).all()
    assert np.allclose(got[fin], want[fin], rtol=rtol, atol=0), np.max(np.abs(got[fin] / want[fin] - 1))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("shape", NONFINITE_SHAPES)
def test_sum_family_and_cumsum_nonfinite(dev, dtype, shape):
    """sum, mean, asum1, asum2, norm2 and cumsum of segments holding +inf,
    -inf, NaN, both infinities or values whose sum overflows: +-inf and NaN
    exactly where NT2's float recurrence has them, every other segment
    finite and accurate (the column, tile and run kernels)."""
    lo, mid, hi = shape
    x3 = _nonfinite_segments(lo, mid, hi, dtype)
    x = x3.reshape(-1)
    rtol = 1e-6 if dtype == np.float32 else 1e-14
    for op in ("sum", "mean", "asum1", "asum2", "norm2"):
        _same_nonfinite(_fold(op, x, lo, mid, hi, dev), _ieee_fold(op, x3, dtype), rtol)
    with np.errstate(all="ignore"):
        want = np.cumsum(x3.astype(np.float64), axis=1).astype(dtype).reshape(-1)
    # numpy's sequential FP64 cumsum is itself only within mid * eps
    _same_nonfinite(_scan("cumsum", x, lo, mid, hi, dev), want, 1e-6 if dtype == np.float32 else 1e-9)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("n", [1000, (1 << 20) + 3, (1 << 24) + 1])
def test_whole_table_sums_nonfinite(dev, dtype, n):
    """Whole-table sum (all three summation modes), dot, the sum family and
    cumsum with one +inf, -inf, NaN, both infinities, or an overflowing
    total: the IEEE value of the plain sum, never a NaN from the
    compensation (single- and two-launch grids, the vector body and the
    scalar tail)."""
    import torch

    sfx = "f32" if dtype == np.float32 else "f64"
    for case in NONFINITE[1:]:
        x3 = np.random.default_rng(7).uniform(0.5, 1.5, (1, n, 1)).astype(dtype)
        if case == "+inf":
            x3[0, n // 3, 0] = np.inf
        elif case == "-inf":
            x3[0, n - 1, 0] = -np.inf  # in the scalar tail when n % 4 != 0
        elif case == "nan":
            x3[0, n // 2, 0] = np.nan
        elif case == "+inf-inf":
            x3[0, 1, 0], x3[0, n - 2, 0] = np.inf, -np.inf
        else:
            x3[:] = np.finfo(dtype).max / 4
        x = x3.reshape(-1)
        g = _dev_tensor(x, dev)
        want = _ieee_fold("sum", x3, dtype)[0]
        for mode in (nt2_amd.SUM_NAIVE, nt2_amd.SUM_KAHAN, nt2_amd.SUM_PAIRWISE):
            if sfx == "f64" and mode != nt2_amd.SUM_KAHAN:
                continue
            _same_nonfinite([ops.sum(g, mode=mode).item()], [want], 1e-6)
        with np.errstate(all="ignore"):
            wd = np.dot(x.astype(np.float64), x.astype(np.float64))
        _same_nonfinite([ops.dot(g, g).item()], [np.float64(dtype(wd)) if np.isfinite(wd) else wd], 1e-6)
        for op in ("sum", "mean", "asum1", "asum2", "norm2"):
            _same_nonfinite(ops.global_fold(op, g).cpu().numpy(), _ieee_fold(op, x3, dtype), 1e-6)
        with np.errstate(all="ignore"):
            wc = np.cumsum(x.astype(np.float64)).astype(dtype)
        _same_nonfinite(_scan("cumsum", x, 1, n, 1, dev), wc, 1e-6 if dtype == np.float32 else 1e-9)
        del g
        torch.cuda.synchronize()


PROD_CASES = ("finite", "+inf", "-inf", "nan", "0*inf", "-0", "overflow", "underflow", "underflow-neg")


def _prod_segments(lo, mid, hi, dtype, seed=41):
    """(hi, mid, lo) data whose segment (h, i) is case PROD_CASES[(h*lo+i) % 9]."""
    rng = np.random.default_rng(seed)
    # (finite segments stay far from the subnormals, where numpy's FP64
    # recurrence would lose the bits the reference needs)
    x = rng.uniform(0.99, 1.01, (hi, mid, lo)).astype(dtype)
    big, small = (dtype(1e20), dtype(1e-20)) if dtype == np.float32 else (dtype(1e200), dtype(1e-200))
    for h in range(hi):
        for i in range(lo):
            case = PROD_CASES[(h * lo + i) % len(PROD_CASES)]
            k = rng.integers(0, mid, 2)
            k1 = (k[0] + 1 + k[1] % (mid - 1)) % mid if mid > 1 else k[0]
            if case == "+inf":
                x[h, k[0], i] = np.inf
            elif case == "-inf":
                x[h, k[0], i] = -np.inf
            elif case == "nan":
                x[h, k[0], i] = np.nan
            elif case == "0*inf":
                x[h, k[0], i], x[h, k1, i] = 0.0, np.inf
            elif case == "-0":
                x[h, k[0], i] = -0.0
            elif case == "overflow" and mid > 2:
                x[h, :, i] = big
            elif case == "underflow" and mid > 3:
                x[h, :, i] = small
            elif case == "underflow-neg" and mid > 3:
                x[h, :, i] = -small
    return x


def _same_ieee(got, want, rtol):
    """NaN, +-inf and signed zeros where want has them; the rest within rtol."""
    got, want = np.asarray(got, np.float64), np.asarray(want, np.float64)
    assert (np.isnan(got) == np.isnan(want)).all(), np.flatnonzero(np.isnan(got) != np.isnan(want))[:5]
    spec = np.isinf(want) | (want == 0)
    assert (got[spec] == want[spec]).all(), (got[spec][got[spec] != want[spec]][:5], want[spec][got[spec] != want[spec]][:5])
    assert (np.signbit(got[spec]) == np.signbit(want[spec])).all()
    fin = np.isfinite(want) & (want != 0)
    assert np.allclose(got[fin], want[fin], rtol=rtol, atol=0), np.max(np.abs(got[fin] / want[fin] - 1))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("shape", NONFINITE_SHAPES + [(1, (1 << 20) + 3, 1), (1, 1 << 18, 2)])
def test_prod_and_cumprod_specials(dev, dtype, shape):
    """prod and cumprod of segments holding +-inf, NaN, 0 with inf, -0, and
    totals that overflow or underflow the element type (the FP64 mantissa
    and separate exponent must end in +-inf and signed zeros exactly where
    the true product does): the IEEE value of the product, rounded once."""
    lo, mid, hi = shape
    x3 = _prod_segments(lo, mid, hi, dtype)
    x = x3.reshape(-1)
    with np.errstate(all="ignore"):
        wp = np.prod(x3.astype(np.float64), axis=1).astype(dtype).reshape(-1)
        wc = np.cumprod(x3.astype(np.float64), axis=1).astype(dtype).reshape(-1)
    rtol = 1e-6 if dtype == np.float32 else 1e-9
    _same_ieee(_fold("prod", x, lo, mid, hi, dev), wp, rtol)
    _same_ieee(_scan("cumprod", x, lo, mid, hi, dev), wc, rtol)


def _fold_scan_at(kind, op, x, lo, mid, hi, dev, offset):
    """nt2h_fold/nt2h_scan with the input (and, for a scan, the output) placed
    `offset` elements into its allocation (offset 1: not 16-byte aligned)."""
    import torch

    sfx = _sfx(x)
    gx = _dev_tensor(np.concatenate([np.zeros(offset, x.dtype), x]), dev)[offset:]
    nout = lo * hi if kind == "fold" else x.size
    y = torch.zeros(nout + offset, dtype=gx.dtype, device=dev)[offset:]
    if kind == "fold":
        nt2_amd.call("nt2h_fold_" + sfx, nt2_amd.FOLD_OPS[op], gx.data_ptr(), lo, mid, hi, y.data_ptr(),
                     torch.cuda.current_stream().cuda_stream)
    else:
        nt2_amd.call("nt2h_scan_" + sfx, nt2_amd.SCAN_OPS[op], gx.data_ptr(), lo, mid, hi, y.data_ptr(),
                     torch.cuda.current_stream().cuda_stream)
    return y.cpu().numpy()


@pytest.mark.parametrize("shape", [(1, 5000, 37), (7, 3000, 5), (16, 5000, 3), (1, (1 << 20) + 3, 1), (4, 30000, 1),
                                   (64, 4097, 1), (1, 1 << 18, 2), (1024, 300, 2), (3, 5, 2)])
def test_folds_and_scans_of_unaligned_tables(dev, ref, shape):
    """Input and output one element past a 16-byte boundary (the element-wise
    loads and stores of every fold and scan path): maximum, minimum, cummax
    and cummin bit-identical to NT2's recurrence, the sums and products and
    cumsum/cumprod within 1e-6 of the FP64 values, as for aligned tables."""
    lo, mid, hi = shape
    x = _specials_data(lo * mid * hi, 51)
    for op in ("maximum", "minimum"):
        got = _fold_scan_at("fold", op, x, lo, mid, hi, dev, 1)
        assert _bits_eq(got, ref.fold_lmh(op, x, lo, mid, hi)).all(), op
    for op in ("cummax", "cummin"):
        got = _fold_scan_at("scan", op, x, lo, mid, hi, dev, 1)
        assert _bits_eq(got, ref.scan_lmh(op, x, lo, mid, hi)).all(), op
    v = np.random.default_rng(52).uniform(0.5, 1.5, lo * mid * hi).astype(np.float32)
    v3 = v.astype(np.float64).reshape(hi, mid, lo)
    p = np.random.default_rng(53).uniform(0.999, 1.001, lo * mid * hi).astype(np.float32)
    p3 = p.astype(np.float64).reshape(hi, mid, lo)
    checks = (("fold", "sum", v, v3.sum(axis=1).reshape(-1)), ("fold", "prod", p, p3.prod(axis=1).reshape(-1)),
              ("scan", "cumsum", v, np.cumsum(v3, axis=1).reshape(-1)),
              ("scan", "cumprod", p, np.cumprod(p3, axis=1).reshape(-1)))
    for kind, op, data, want in checks:
        for offset in (0, 1):
            got = _fold_scan_at(kind, op, data, lo, mid, hi, dev, offset).astype(np.float64)
            assert np.allclose(got, want, rtol=1e-6, atol=0), (kind, op, offset, np.max(np.abs(got / want - 1)))
