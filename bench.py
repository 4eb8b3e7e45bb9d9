#!/usr/bin/env python3
"""Benchmark: Black-Scholes forward + adjoint on one MI355X.

Workload (the headline configuration): 2^24 float options, S~U[5,30],
K~U[1,100], T~U[0.25,10], r=0.02, sigma=0.30 (stored as full streams),
cotangent w~U[-1,1].  One step = one nt2h_bs_fwd_adj_f32 launch over all
options: 6 input streams read, 6 output streams (price, dS, dK, dT, dr,
dsig) written, plus the compensated totals sum(dr), sum(dsig) — 48 bytes of
algorithmic HBM traffic per option.  Inputs are resident in HBM before the
timed region.

Prints one JSON line: options/s (value), the roofline fraction of the
kernel measured with HIP events on its stream, the same tree in plain FP64
on the GPU (speed and the accuracy gap), a device copy of equal traffic,
and the oracle's multithreaded CPU evaluation on a bounded sample
(cpu_baseline).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0  # MI355X HBM3E, published
N_OPTIONS = 1 << 24
BYTES_PER_OPTION = 48  # in: S K T r sig w, out: price dS dK dT dr dsig (float)
WARMUP_MIN_S = 0.5  # minimum duration of the warm-up launches before the timed steps


def _traffic_from_profile():
    """HBM bytes per launch measured with rocprofv3 PMC counters (profiles/)."""
    p = os.path.join(ROOT, "profiles", "bs_fwd_adj_f32_2p24_pmc.txt")
    if not os.path.exists(p):
        return None
    for line in open(p):
        if line.startswith("traffic_bytes_per_launch:"):
            try:
                return int(float(line.split(":", 1)[1].strip()))
            except ValueError:
                return None
    return None


def _cpu_baseline(inputs_np, budget_s=10.0):
    """The oracle's NT2 functors restated branch-free and vectorised
    (#pragma omp simd in AVX-512 / AVX2 / baseline clones picked by the host
    CPU; bit-identical to the scalar restatement) under NT2's
    OpenMP grain policy, on the host cores, over a bounded sample: median of
    the timed passes after one warm-up.  The scalar restatement is timed
    briefly beside it for reference."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import nt2ref

    lib = nt2ref.lib()

    def median_rate(fn, n, budget, max_reps):
        ins = [np.ascontiguousarray(a[:n]) for a in inputs_np]
        outs = [np.empty(n, np.float32) for _ in range(6)]
        fn(ins, outs)  # warm-up, first touch
        ts, t_all = [], time.perf_counter()
        while (time.perf_counter() - t_all < budget or len(ts) < 5) and len(ts) < max_reps:
            t0 = time.perf_counter()
            fn(ins, outs)
            ts.append(time.perf_counter() - t0)
        return n / float(np.median(ts)), len(ts), time.perf_counter() - t_all

    n = len(inputs_np[0])  # the whole workload: 805 MB of streams, past any cache
    rate, reps, el = median_rate(nt2ref.cpu_simd_bs_fwd_adj_f32, n, budget_s, 2000)
    scalar_rate, _, _ = median_rate(nt2ref.cpu_bs_fwd_adj_f32, 1 << 22, 2.0, 200)
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        pass
    return {
        "value": rate,
        "unit": "options/s",
        "cores": int(lib.nt2ref_max_threads()),
        "kind": "port",
        "sample": f"the whole 2^24-option workload; median of {reps} passes in {el:.1f} s "
                  f"(NT2 functors restated branch-free, vectorised with omp simd in AVX-512, AVX2 and baseline clones picked by the host CPU, OpenMP static grains of L2/4 floats)",
        "gbps": BYTES_PER_OPTION * rate / 1e9,
        "cpu_model": model,
        "isa": nt2ref.cpu_simd_isa(),
        "scalar_port_options_per_s": scalar_rate,  # first 2^22 options, median over ~2 s
    }


def fd_residuals(g64, f32_out, outs64, m=1 << 16):
    """Gradient check of the headline kernel on its first m options: central
    differences (h = 1e-3 |x|, in FP64) of the FP64 call price for S, K, T, r
    and sigma, times w, against the float kernel's greeks and the FP64
    kernel's (the latter shows the differences' own truncation error); each
    as max |g - w dP/dx| over the greek's natural scale (|w| x 1, e^{-rT}, S,
    S T, S sqrt T), the scales of the stated bounds."""
    import torch

    import nt2_amd
    from nt2_amd import ops

    x = [t[:m].clone() for t in g64[:5]]
    w = g64[5][:m]
    S, K, T, r = x[0], x[1], x[2], x[3]
    aw = w.abs()
    scales = [aw, aw * torch.exp(-r * T), aw * S, aw * S * T, aw * S * torch.sqrt(T)]
    out = {"options": m, "h": "1e-3 |x|, FP64 central differences", "f32_vs_fd": {}, "fp64_vs_fd": {}}
    for j, nm in enumerate(("dS", "dK", "dT", "dr", "dsig")):
        h = 1e-3 * x[j].abs()
        up = [t.clone() for t in x]
        dn = [t.clone() for t in x]
        up[j] += h
        dn[j] -= h
        pu = ops.bs_fwd(*up, precision=nt2_amd.EXACT)[0]
        pd = ops.bs_fwd(*dn, precision=nt2_amd.EXACT)[0]
        g_fd = w * (pu - pd) / (up[j] - dn[j])
        sc = torch.clamp(scales[j], min=1e-30)
        out["f32_vs_fd"][nm] = float(((f32_out[j + 1][:m].double() - g_fd).abs() / sc).max())
        out["fp64_vs_fd"][nm] = float(((outs64[j + 1][:m] - g_fd).abs() / sc).max())
    return out


def _events(stream, fn, reps):
    """Seconds per launch over `reps` back-to-back launches, after >= 30 ms of
    warm-up launches (clocks ramp up after the host-side input generation)."""
    import torch

    t0 = time.perf_counter()
    while time.perf_counter() - t0 < 0.03:
        fn()
        torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(stream)
    for _ in range(reps):
        fn()
    b.record(stream)
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps / 1e3  # seconds per launch


def _ulp_vs_fp64(y32, y64):
    """Float ulps of y32 against y64 where y64 is a normal float (|y64| >= 2^-126);
    NT2's functors flush below that (exp builds 2^k from exponent bits), so the
    subnormal band is reported apart as a count and a maximum absolute error."""
    import torch

    tiny = torch.finfo(torch.float32).tiny
    normal = (y64.abs() >= tiny) | (y64 == 0)
    both_nan = torch.isnan(y32) & torch.isnan(y64)
    diff = (y32.double() - y64).abs()
    same = (y32.double() == y64) | both_nan
    _, e = torch.frexp(torch.clamp(y64.abs(), min=tiny))
    sp = torch.ldexp(torch.ones_like(y64), (e - 24).to(torch.int32))
    u = torch.where(same | ~normal, torch.zeros_like(diff), diff / sp)
    rel = torch.where(same | ~normal, torch.zeros_like(diff), diff / torch.clamp(y64.abs(), min=1e-300))
    sub = ~normal
    sub_abs = float(torch.where(sub, diff, torch.zeros_like(diff)).max()) if bool(sub.any()) else 0.0
    return float(u.max()), int(u.argmax()), float(rel.max()), int(sub.sum()), sub_abs


def _rounded_ulps(y32, y64):
    """Max std-ULP (IEEE spacing steps) and max NT2-ULP (modules/test/unit/src/ulp.cpp:18-32:
    frexp-aligned difference over float epsilon) of y32 against y64 rounded to float,
    over the elements where the rounded value is a normal float or zero."""
    import torch

    r = y64.float()
    tiny = torch.finfo(torch.float32).tiny
    ok = ((r.abs() >= tiny) | (r == 0)) & torch.isfinite(r) & torch.isfinite(y32)
    ok &= (y32.abs() >= tiny) | (y32 == 0)

    def ordv(t):
        i = t.view(torch.int32).to(torch.int64)
        return torch.where(i < 0, -(1 << 31) - i, i)

    std = torch.where(ok, (ordv(y32) - ordv(r)).abs(), torch.zeros((), dtype=torch.int64, device=r.device))
    m1, e1 = torch.frexp(y32)
    m2, e2 = torch.frexp(r)
    ex = -torch.maximum(e1, e2)
    d = torch.where(e1 == e2, (m1 - m2).abs(), (torch.ldexp(y32, ex) - torch.ldexp(r, ex)).abs())
    nt2 = torch.where(ok & (y32 != r), d / torch.finfo(torch.float32).eps, torch.zeros_like(d))
    return int(std.max()), float(nt2.max())


def report_sweep(dev, stream, n=1 << 28, fns=("exp", "log", "sqrt", "erfc", "tanh", "sin", "cos")):
    """Precision sweep: each functor over 2^28 floats, FP32 (EXACT, FUSED) vs FP64 on the GPU."""
    import torch

    from nt2_amd import EXACT, FUSED, SUM_KAHAN, SUM_NAIVE, SUM_PAIRWISE, ops, synth
    from nt2_amd.layout import streams

    lines = []
    for fn in fns:
        x, y = streams(2, n, torch.float32, dev)
        x.copy_(torch.from_numpy(synth.sweep_args(fn, n)))
        x64, y64 = streams(2, n, torch.float64, dev)
        x64.copy_(x)
        ops.map(fn, x64, EXACT, out=y64)
        t64 = _events(stream, lambda: ops.map(fn, x64, EXACT, out=y64), 3)
        rec = {"report": "precision_sweep", "functor": fn, "n": n, "args": synth.SWEEP[fn],
               "f64_us": t64 * 1e6, "f64_gbps": 16 * n / t64 / 1e9}
        for name, pr in (("exact", EXACT), ("fused", FUSED)):
            ops.map(fn, x, pr, out=y)
            t = _events(stream, lambda: ops.map(fn, x, pr, out=y), 5)
            u, am, rel, nsub, sub_abs = _ulp_vs_fp64(y, y64)
            std_u, nt2_u = _rounded_ulps(y, y64)
            rec[name] = {"us": t * 1e6, "gbps": 8 * n / t / 1e9, "hbm_frac": 8 * n / t / 8e12,
                         "max_ulp_vs_fp64": u, "worst_arg": float(x[am]), "max_rel_vs_fp64": rel,
                         "max_std_ulp_vs_fp64_rounded": std_u, "max_nt2_ulp_vs_fp64_rounded": nt2_u,
                         "subnormal_band_count": nsub, "subnormal_band_max_abs_err": sub_abs,
                         "speedup_vs_f64": t64 / t}
        lines.append(rec)
        del x, y, x64, y64
        torch.cuda.empty_cache()
    a_np, b_np = synth.sumdot_inputs(n)
    a, b = streams(2, n, torch.float32, dev)
    a.copy_(torch.from_numpy(a_np))
    b.copy_(torch.from_numpy(b_np))
    a64, b64 = streams(2, n, torch.float64, dev)
    a64.copy_(a)
    b64.copy_(b)
    s64 = ops.sum(a64).item()
    d64 = ops.dot(a64, b64).item()
    t64s = _events(stream, lambda: ops.sum(a64), 5)
    t64d = _events(stream, lambda: ops.dot(a64, b64), 5)
    # NT2's scalar inner_fold order (left to right in float), for information
    seq_s = float(np.cumsum(a_np, dtype=np.float32)[-1])
    seq_d = float(np.cumsum(a_np * b_np, dtype=np.float32)[-1])
    rec = {"report": "sum_dot", "n": n, "values": "N(1, 1e-3)", "f64": {"sum": s64, "dot": d64,
           "sum_us": t64s * 1e6, "dot_us": t64d * 1e6},
           "nt2_sequential_order": {"sum_ulp_f32": abs(seq_s - s64) / float(np.spacing(np.float32(s64))),
                                    "dot_ulp_f32": abs(seq_d - d64) / float(np.spacing(np.float32(d64)))}}
    for name, mode in (("naive", SUM_NAIVE), ("kahan", SUM_KAHAN), ("pairwise", SUM_PAIRWISE)):
        s = ops.sum(a, mode).item()
        d = ops.dot(a, b, mode).item()
        ts = _events(stream, lambda: ops.sum(a, mode), 5)
        td = _events(stream, lambda: ops.dot(a, b, mode), 5)
        ulp_s = abs(float(np.float32(s)) - s64) / float(np.spacing(np.float32(s64)))
        ulp_d = abs(float(np.float32(d)) - d64) / float(np.spacing(np.float32(d64)))
        rec[name] = {"sum_rel_err": abs(s - s64) / abs(s64), "sum_ulp_f32": ulp_s, "sum_us": ts * 1e6,
                     "sum_gbps": 4 * n / ts / 1e9, "dot_rel_err": abs(d - d64) / abs(d64), "dot_ulp_f32": ulp_d,
                     "dot_us": td * 1e6, "dot_gbps": 8 * n / td / 1e9}
    lines.append(rec)
    return lines


def report_bs_fwd(dev, stream, n=1 << 26):
    """Black-Scholes forward (call, put) over 2^26 options: FP32 EXACT / FUSED vs FP64."""
    import torch

    from nt2_amd import EXACT, FUSED, ops, synth
    from nt2_amd.layout import streams

    d = synth.bs_inputs(n)
    g = streams(5, n, torch.float32, dev)
    for t, k in zip(g, ("S", "K", "T", "r", "sig")):
        t.copy_(torch.from_numpy(d[k]))
    g64 = streams(5, n, torch.float64, dev)
    for t64, t in zip(g64, g):
        t64.copy_(t)
    c64, p64 = ops.bs_fwd(*g64, precision=EXACT)
    t64 = _events(stream, lambda: ops.bs_fwd(*g64, precision=EXACT), 5)
    rec = {"report": "bs_fwd", "options": n, "f64_us": t64 * 1e6, "f64_gbps": 56 * n / t64 / 1e9}
    S = g64[0]
    for name, pr in (("exact", EXACT), ("fused", FUSED)):
        c, p = ops.bs_fwd(*g, precision=pr)
        t = _events(stream, lambda: ops.bs_fwd(*g, precision=pr), 10)
        sc = torch.maximum(c64.abs(), 1e-6 * S)
        sp_ = torch.maximum(p64.abs(), 1e-6 * S)
        rec[name] = {"us": t * 1e6, "options_per_s": n / t, "gbps": 28 * n / t / 1e9, "hbm_frac": 28 * n / t / 8e12,
                     "call_scaled_err": float(((c.double() - c64).abs() / sc).max()),
                     "put_scaled_err": float(((p.double() - p64).abs() / sp_).max()),
                     "speedup_vs_f64": t64 / t}
    return [rec]


def report_saxpy(dev, stream, n=1 << 20):
    import torch

    from nt2_amd import SUM_KAHAN, ops, synth

    from nt2_amd import call

    x_np, y_np = synth.saxpy_inputs(n)
    x, y = torch.from_numpy(x_np).to(dev), torch.from_numpy(y_np).to(dev)
    t = _events(stream, lambda: ops.saxpy_(2.5, x, y), 50)
    td = _events(stream, lambda: ops.dot(x, y, SUM_KAHAN), 50)
    # GPU time per call without the host's submission rate: 100 calls
    # recorded in one HIP graph, replayed
    out = torch.zeros(1, dtype=torch.float64, device=dev)

    def saxpy():
        call("nt2h_saxpy_f32", 2.5, x.data_ptr(), y.data_ptr(), n, torch.cuda.current_stream().cuda_stream)

    def sdot():
        call("nt2h_dot_f32", x.data_ptr(), y.data_ptr(), n, SUM_KAHAN, out.data_ptr(),
             torch.cuda.current_stream().cuda_stream)

    tg, tdg = _graph_seconds(saxpy), _graph_seconds(sdot)
    return [{"report": "saxpy_sdot", "n": n, "saxpy_us": t * 1e6, "saxpy_gbps": 12 * n / t / 1e9,
             "sdot_us": td * 1e6, "sdot_gbps": 8 * n / td / 1e9,
             "graph_saxpy_us": tg * 1e6, "graph_saxpy_gbps": 12 * n / tg / 1e9,
             "graph_sdot_us": tdg * 1e6, "graph_sdot_gbps": 8 * n / tdg / 1e9,
             "note": "2^20 elements: launch-latency bound; *_us are back-to-back calls from Python, "
                     "graph_* the same calls replayed from one HIP graph of 100"}]


def _graph_seconds(fn, calls=100, replays=10):
    """Seconds per call with `calls` calls recorded in one HIP graph and the
    graph replayed `replays` times (events around the replays)."""
    import torch

    cs = torch.cuda.Stream()  # the library keeps workspaces per stream:
    cs.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(cs):  # allocate them on the capture stream first
        fn()
    cs.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=cs):
        for _ in range(calls):
            fn()
    g.replay()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(replays):
        g.replay()
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / (calls * replays) / 1e3


def report_placement(dev, stream, n=1 << 24):
    """The headline kernel on coloured tables vs 12 separate allocations."""
    import torch

    from nt2_amd import ops, synth
    from nt2_amd.layout import streams

    d = synth.bs_inputs(n, with_w=True)
    host = [torch.from_numpy(d[k]) for k in ("S", "K", "T", "r", "sig", "w")]
    col = streams(12, n, torch.float32, dev)
    for t, h in zip(col, host):
        t.copy_(h)
    sep = [h.to(dev) for h in host] + [torch.empty(n, dtype=torch.float32, device=dev) for _ in range(6)]
    rec = {"report": "placement", "options": n}
    for name, tabs in (("coloured", col), ("separate_allocations", sep)):
        offs = sorted((t.data_ptr() - tabs[0].data_ptr()) % (2 << 20) for t in tabs)
        t = _events(stream, lambda: ops.bs_fwd_adj(*tabs[:6], totals=True, outs=tabs[6:]), 50)
        rec[name] = {"kernel_us": t * 1e6, "gbps": 48 * n / t / 1e9, "offsets_mod_2MiB_KiB": [o // 1024 for o in offs]}
    return [rec]


def report_folds(dev, stream):
    """The other reduction functors and the scans: whole-table over 2^28
    floats, along each dimension of a 16384 x 8192 table, along dimension 2
    of a 64 x 2^21 and a 4 x 2^26 table (few outputs) and of a 256 x 4096 x
    256 table
    (decimal GB/s of algorithmic traffic: 4 B/element for a fold, 8
    B/element for a scan)."""
    import torch

    from nt2_amd import FOLD_OPS, SCAN_OPS, call
    from nt2_amd.layout import streams

    lines = []
    st = torch.cuda.current_stream().cuda_stream
    n = 1 << 28
    x, y = streams(2, n, torch.float32, dev)
    x.normal_()
    cases = [("global 2^28", 1, n, 1), ("dim1 16384x8192", 1, 16384, 8192), ("dim2 16384x8192", 16384, 8192, 1),
             ("dim2 64x2^21", 64, 1 << 21, 1), ("dim2 4x2^26", 4, 1 << 26, 1),
             ("dim2 256x4096x256", 256, 4096, 256)]
    for name, lo, mid, hi in cases:
        rec = {"report": "folds", "shape": name, "lo_mid_hi": [lo, mid, hi]}
        out = torch.empty(lo * hi, dtype=torch.float32, device=dev)
        for op in ("sum", "maximum", "prod", "norm2", "mean"):
            t = _events(stream, lambda: call("nt2h_fold_f32", FOLD_OPS[op], x.data_ptr(), lo, mid, hi,
                                             out.data_ptr(), st), 10)
            rec[op] = {"us": t * 1e6, "gbps": 4 * lo * mid * hi / t / 1e9}
        for op in ("cumsum", "cummax", "cumprod"):
            t = _events(stream, lambda: call("nt2h_scan_f32", SCAN_OPS[op], x.data_ptr(), lo, mid, hi,
                                             y.data_ptr(), st), 10)
            rec[op] = {"us": t * 1e6, "gbps": 8 * lo * mid * hi / t / 1e9}
        lines.append(rec)
    del x, y
    return lines


def report_scans(dev, stream):
    """Scans of medium runs, where launch count matters: runs of 2^14 to 2^20
    floats (up to 512 tiles of 4096 or 8192 floats in all) take one launch
    with decoupled look-back, larger ones three passes (decimal GB/s of 8
    B/element; times per call replayed from a HIP graph, and issued one by one
    from Python for comparison)."""
    import torch

    from nt2_amd import SCAN_OPS, call
    from nt2_amd.layout import streams

    lines = []
    n = 1 << 23
    x, y = streams(2, n, torch.float32, dev)
    x.normal_()
    for lo, mid, hi in ((1, 1 << 14, 1), (1, 1 << 16, 1), (1, 1 << 18, 1), (1, 1 << 19, 1), (1, 1 << 20, 1),
                        (1, 1 << 16, 16), (1, 1 << 16, 64), (1, 1 << 21, 1), (1, 1 << 16, 128)):
        t4 = -(-mid // 4096)  # tiles of 4096 floats per run; 8192-float tiles past 64 per run or 256 in all
        tiles = t4 if (t4 <= 64 and t4 * hi <= 256) else -(-mid // 8192)
        path = "look-back" if tiles <= 128 and tiles * hi <= 512 else "three passes"
        rec = {"report": "scans", "lo_mid_hi": [lo, mid, hi], "path": path}  # cumsum and cummax
        for op in ("cumsum", "cummax", "cumprod"):
            def scan():
                call("nt2h_scan_f32", SCAN_OPS[op], x.data_ptr(), lo, mid, hi, y.data_ptr(),
                     torch.cuda.current_stream().cuda_stream)
            t = _graph_seconds(scan)  # GPU time per call: 100 calls replayed from one HIP graph
            tp = _events(stream, scan, 200)  # the same calls issued one by one from Python
            rec[op] = {"us": t * 1e6, "gbps": 8 * lo * mid * hi / t / 1e9, "python_calls_us": tp * 1e6}
        lines.append(rec)
    del x, y
    return lines


def report_host(dev, stream, n=1 << 24):
    """Host tables: staged pageable copies (nt2h_upload / nt2h_download) against
    hipMemcpy from pageable and from pinned memory, and the Black-Scholes
    fwd+adj over host tables (PCIe-inclusive options/s, blocks of 2^20
    options overlapped on two streams)."""
    import torch

    from nt2_amd import call, synth

    st = torch.cuda.current_stream().cuda_stream
    nbytes = 256 << 20
    host = np.ones(nbytes, np.uint8)
    back = np.zeros(nbytes, np.uint8)
    d = torch.empty(nbytes, dtype=torch.uint8, device=dev)
    pinned = torch.ones(nbytes, dtype=torch.uint8).pin_memory()
    th = torch.from_numpy(host)

    def wall(fn, reps=4):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps

    rec = {"report": "host", "bytes": nbytes}
    rec["upload_staged_gbps"] = nbytes / wall(lambda: call("nt2h_upload", d.data_ptr(), host.ctypes.data, nbytes, st)) / 1e9
    rec["upload_hipmemcpy_pageable_gbps"] = nbytes / wall(lambda: d.copy_(th)) / 1e9
    rec["upload_pinned_gbps"] = nbytes / wall(lambda: d.copy_(pinned, non_blocking=True)) / 1e9
    rec["download_staged_gbps"] = nbytes / wall(lambda: call("nt2h_download", back.ctypes.data, d.data_ptr(), nbytes, st)) / 1e9
    rec["download_hipmemcpy_pageable_gbps"] = nbytes / wall(lambda: th.copy_(d)) / 1e9
    rec["download_pinned_gbps"] = nbytes / wall(lambda: pinned.copy_(d, non_blocking=True)) / 1e9
    del d, pinned
    dd = synth.bs_inputs(n, with_w=True)
    ins = [dd[k] for k in ("S", "K", "T", "r", "sig", "w")]
    outs = [np.empty(n, np.float32) for _ in range(6)]
    tot = np.zeros(2)
    args = [a.ctypes.data for a in ins] + [o.ctypes.data for o in outs] + [tot.ctypes.data, n, 1 << 20]
    t = wall(lambda: call("nt2h_bs_fwd_adj_f32_host", *args), 3)
    rec["bs_fwd_adj_host"] = {"options": n, "block": 1 << 20, "s": t, "options_per_s": n / t,
                              "pcie_gbps": 48 * n / t / 1e9}
    return [rec]


def report_colfold(dev, stream, rows=16384, cols=8192):
    """Column fold + adjoint over 16384 x 8192, FAST and FUSED arithmetic: time,
    and error vs a torch FP64 evaluation on 64 columns (over each output's sum of
    |terms|)."""
    import torch

    import nt2_amd
    from nt2_amd import ops, synth

    X_np, s_np, k_np, p_np = synth.colfold_inputs(rows, cols)
    X = torch.from_numpy(X_np).to(dev)
    s, k, p = (torch.from_numpy(a).to(dev) for a in (s_np, k_np, p_np))
    sel = torch.arange(0, cols, cols // 64, device=dev)
    Xd, sd, kd, pd = X[sel].double(), s[sel].double()[:, None], k[sel].double()[:, None], p[sel].double()[:, None]
    e = torch.exp(-Xd * Xd / (2 * sd * sd))
    arg = kd * Xd + pd
    zc, sn = e * torch.cos(arg), e * torch.sin(arg)
    ref = [zc.sum(1), (zc * Xd * Xd).sum(1) / sd[:, 0] ** 3, -(sn * Xd).sum(1), -sn.sum(1)]
    mag = [e.sum(1), (e * Xd * Xd).sum(1) / sd[:, 0] ** 3, (e * Xd.abs()).sum(1), e.sum(1)]
    # central differences (h = 1e-3 |param|) of the FP64 fold for s, k, p
    def zfold(sv, kv, pv):
        return (torch.exp(-Xd * Xd / (2 * sv * sv)) * torch.cos(kv * Xd + pv)).sum(1)

    fd = []
    for j in range(3):
        prm = [sd, kd, pd]
        h = 1e-3 * prm[j].abs()
        up, dn = list(prm), list(prm)
        up[j], dn[j] = prm[j] + h, prm[j] - h
        fd.append((zfold(*up) - zfold(*dn)) / (up[j] - dn[j])[:, 0])
    recs = []
    for name, prec in (("fast", nt2_amd.FAST), ("fused", nt2_amd.FUSED)):
        z, ds, dk, dp = ops.colfold_gc(X, s, k, p, precision=prec)
        t = _events(stream, lambda: ops.colfold_gc(X, s, k, p, precision=prec), 20)
        tf = _events(stream, lambda: ops.colfold_gc(X, s, k, p, grad=False, precision=prec), 20)
        errs = [float(((g[sel].double() - r).abs() / m).max()) for g, r, m in zip((z, ds, dk, dp), ref, mag)]
        recs.append({"report": "colfold", "precision": name, "rows": rows, "cols": cols,
                     "fwd_adj_us": t * 1e6, "fwd_only_us": tf * 1e6,
                     "fwd_adj_gbps": 4 * rows * cols / t / 1e9, "hbm_frac": 4 * rows * cols / t / 8e12,
                     "fwd_only_hbm_frac": 4 * rows * cols / tf / 8e12, "elements_per_s": rows * cols / t,
                     "err_vs_fp64_over_sum_abs_terms": dict(zip(("z", "dz_ds", "dz_dk", "dz_dp"), errs)),
                     # gradient check: against FP64 central differences, and the FP64
                     # analytic gradients against the same (the differences' truncation)
                     "fd_residual_over_sum_abs_terms": {
                         nm: float(((g[sel].double() - f).abs() / mm).max())
                         for nm, g, f, mm in zip(("dz_ds", "dz_dk", "dz_dp"), (ds, dk, dp), fd, mag[1:])},
                     "fp64_vs_fd_over_sum_abs_terms": {
                         nm: float(((a - f).abs() / mm).max())
                         for nm, a, f, mm in zip(("dz_ds", "dz_dk", "dz_dp"), ref[1:], fd, mag[1:])}})
    return recs


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--report", default="", help="extra report lines after the headline: "
                    "comma list of placement,sweep,bs_fwd,saxpy,colfold,folds,scans,host (or all)")
    ap.add_argument("--sweep-fns", default="exp,log,sqrt,erfc,tanh,sin,cos",
                    help="functors of the precision sweep report")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: bench.py measures one GPU only (asked for --gpus {args.gpus}); "
              "the evaluator shards by option range with no exchange except the two FP64 totals",
              file=sys.stderr)
        sys.exit(2)

    import torch

    import nt2_amd
    from nt2_amd import ops, synth
    from nt2_amd.layout import streams

    nt2_amd.call("nt2h_init", 0)
    dev = torch.device("cuda:0")
    n = N_OPTIONS
    d = synth.bs_inputs(n, with_w=True)
    keys = ("S", "K", "T", "r", "sig", "w")
    host = [d[k] for k in keys]
    # the 12 tables of the loop nest, placed 4 KiB apart modulo 2 MiB
    tabs = streams(12, n, torch.float32, dev)
    g, outs = tabs[:6], tabs[6:]
    for t, a in zip(g, host):
        t.copy_(torch.from_numpy(a))
    tot = torch.empty(2, dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream()

    def step():
        return ops.bs_fwd_adj(*g, totals=True, outs=outs, tot_out=tot)

    # W warm-up steps, continued until at least WARMUP_MIN_S of back-to-back
    # launches have run: after an idle period (input generation, a fresh box)
    # the first ~100 launches run ~5% slow while the clocks ramp up
    warm_launches, tw = 0, time.perf_counter()
    while True:
        batch = args.warmup if warm_launches == 0 and args.warmup > 0 else 50
        for _ in range(batch):
            step()
        warm_launches += batch
        torch.cuda.synchronize()
        if time.perf_counter() - tw >= WARMUP_MIN_S:
            break
    warm_s = time.perf_counter() - tw

    # K back-to-back launches bracketed by synchronisation; HIP events on the
    # launch stream give the kernel time per launch (gaps included)
    e_start, e_end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e_start.record(stream)
    for i in range(args.steps):
        res = step()
    e_end.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    avg_kern_s = e_start.elapsed_time(e_end) / 1e3 / args.steps
    ms_per_step = wall / args.steps * 1e3
    value = n * args.steps / wall
    achieved = BYTES_PER_OPTION * n / avg_kern_s / 1e9
    totals32 = res[6].cpu().numpy()
    f32_out = [o.clone() for o in outs]
    # per-launch event pairs: median of 50 launches
    per = []
    for _ in range(50):
        p0, p1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        p0.record(stream)
        step()
        p1.record(stream)
        per.append((p0, p1))
    torch.cuda.synchronize()
    med_us = float(np.median([a.elapsed_time(b) * 1e3 for a, b in per]))

    # plain FP64 evaluation of the same tree on the same (promoted) inputs
    tabs64 = streams(12, n, torch.float64, dev)
    g64, outs64 = tabs64[:6], tabs64[6:]
    for t64, t in zip(g64, g):
        t64.copy_(t)
    for _ in range(5):
        r64 = ops.bs_fwd_adj(*g64, totals=True, outs=outs64)
    torch.cuda.synchronize()
    steps64 = max(10, args.steps // 4)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps64):
        r64 = ops.bs_fwd_adj(*g64, totals=True, outs=outs64)
    e1.record(stream)
    torch.cuda.synchronize()
    t64 = e0.elapsed_time(e1) / 1e3 / steps64
    totals64 = r64[6].cpu().numpy()

    # accuracy of the float kernel against the FP64 evaluation
    S64 = g64[0]
    acc = {}
    names = ["price", "dS", "dK", "dT", "dr", "dsig"]
    w64 = g64[5].abs()
    scales = [torch.maximum(outs64[0].abs(), 1e-6 * S64), w64, w64 * torch.exp(-g64[3] * g64[2]), w64 * S64,
              w64 * S64 * g64[2], w64 * S64 * torch.sqrt(g64[2])]
    for i, nm in enumerate(names):
        diff = (f32_out[i].double() - outs64[i]).abs()
        acc[nm] = float((diff / torch.clamp(scales[i], min=1e-30)).max())
    # price in float ULPs where the FP64 price is a normal float
    p64 = outs64[0]
    sp = torch.clamp(p64.abs().float(), min=torch.finfo(torch.float32).tiny)
    ulp = torch.nextafter(sp, torch.full_like(sp, float("inf"))) - sp
    price_ulp = (f32_out[0].double() - p64).abs() / ulp.double()
    big = p64.abs() >= 1e-3 * S64
    acc["price_max_ulp_where_price>=1e-3*S"] = float(price_ulp[big].max())
    acc["totals_rel_err"] = [float(abs(totals32[i] - totals64[i]) / max(abs(totals64[i]), 1e-300)) for i in range(2)]
    # the stated bounds (nt2_amd.accuracy), evaluated on this run's outputs
    from nt2_amd import accuracy

    m = accuracy.metrics(tuple(g64[:6]), f32_out, outs64, totals32, totals64)
    bound = {"measured": m, "bound": accuracy.BOUNDS, "violations": accuracy.check(m)}
    bound["pass"] = not bound["violations"]
    fd = fd_residuals(g64, f32_out, outs64)

    # device copy of the same byte count (read 24 B + write 24 B per option)
    src = torch.empty(n * 6, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    c0, c1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    c0.record(stream)
    for _ in range(20):
        dst.copy_(src)
    c1.record(stream)
    torch.cuda.synchronize()
    copy_gbps = 2 * src.numel() * 4 / (c0.elapsed_time(c1) / 1e3 / 20) / 1e9
    del src, dst

    # saxpy over 2^26 floats (8 B read + 4 B written per element): the
    # calibration kernel for the PMC byte counters
    ns = 1 << 26
    xs = torch.ones(ns, dtype=torch.float32, device=dev)
    ys = torch.zeros(ns, dtype=torch.float32, device=dev)
    ops.saxpy_(0.5, xs, ys)
    torch.cuda.synchronize()
    s0, s1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s0.record(stream)
    for _ in range(10):
        ops.saxpy_(0.5, xs, ys)
    s1.record(stream)
    torch.cuda.synchronize()
    saxpy_gbps = 12 * ns / (s0.elapsed_time(s1) / 1e3 / 10) / 1e9
    del xs, ys

    traffic = _traffic_from_profile()
    line = {
        "metric": "BS fwd+adjoint options/s & % HBM peak; FP32-compensated vs FP64 at stated ULP",
        "value": value,
        "unit": "options/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "warmup_min_s": WARMUP_MIN_S,
        "warmup_launches": warm_launches,
        "warmup_s": warm_s,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded SplitMix64 streams of the named distributions)",
        "config": {
            "workload": "Black-Scholes forward + adjoint, 2^24 float options, one fused kernel",
            "options": n,
            "inputs": "S~U[5,30] K~U[1,100] T~U[0.25,10] r=0.02 sigma=0.30 w~U[-1,1], seed 0",
            "outputs": "price, dS, dK, dT, dr, dsig (w-weighted), totals sum(dr) sum(dsig) in FP64",
            "bytes_per_option": BYTES_PER_OPTION,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBPS,
            "traffic": traffic,
            "traffic_source": "profile-derived, not measured in this run: rocprofv3 --pmc FETCH_SIZE and "
                              "WRITE_SIZE of this kernel, scaled by the saxpy calibration "
                              "(profiles/bs_fwd_adj_f32_2p24_pmc.txt)",
        },
        "kernel_us": avg_kern_s * 1e6,
        "kernel_us_median_of_50_event_pairs": med_us,
        "hbm_pct_of_peak": 100.0 * achieved / HBM_PEAK_GBPS,
        "placement": "12 tables in one allocation, each on its own 4 KiB colour (nt2_amd.layout.streams)",
        "copy_gbps_same_bytes": copy_gbps,
        "saxpy_2p26_gbps": saxpy_gbps,
        "fp64": {
            "options_per_s": n / t64,
            "kernel_us": t64 * 1e6,
            "achieved_gbps": 2 * BYTES_PER_OPTION * n / t64 / 1e9,
            "fp32_speedup": t64 / avg_kern_s,
        },
        "accuracy_vs_fp64": acc,
        "accuracy_bound": bound,
        "gradient_check": fd,
    }
    if not args.no_cpu:
        line["cpu_baseline"] = _cpu_baseline(host)
        line["cpu_baseline"]["gpu_over_cpu"] = value / line["cpu_baseline"]["value"]
    print(json.dumps(line), flush=True)
    want = set(args.report.split(",")) if args.report else set()
    if "all" in want:
        want = {"saxpy", "bs_fwd", "sweep", "colfold", "placement", "folds", "scans", "host"}
    del g, outs, tabs, g64, outs64, tabs64
    torch.cuda.empty_cache()
    for name, fn in (("placement", report_placement), ("saxpy", report_saxpy), ("bs_fwd", report_bs_fwd),
                     ("colfold", report_colfold), ("folds", report_folds), ("scans", report_scans),
                     ("host", report_host),
                     ("sweep", report_sweep)):
        if name in want:
            recs = fn(dev, stream, fns=tuple(args.sweep_fns.split(","))) if name == "sweep" else fn(dev, stream)
            for rec in recs:
                print(json.dumps(rec), flush=True)
            torch.cuda.empty_cache()


if __name__ == "__main__":
    main()
