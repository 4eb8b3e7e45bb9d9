"""The allocation-free Black-Scholes forward + adjoint entry point
(nt2h_bs_fwd_adj_f32_noalloc) against nt2h_bs_fwd_adj_f32, through the C ABI:
bit-identical outputs and totals on two calls in a row (the second smaller),
its size limit, its refusal of host totals, and replays of a recorded graph."""
import ctypes as C

import numpy as np
import pytest

import nt2_amd
from nt2_amd import synth

pytestmark = pytest.mark.gpu

MAX_OPTIONS = 1 << 24


def _tables(n, dev, torch):
    d = synth.bs_inputs(n, with_w=True)
    ins = [torch.from_numpy(d[k]).to(dev) for k in ("S", "K", "T", "r", "sig", "w")]
    return ins


def _call(fn, ins, outs, tot, n, stream):
    L = nt2_amd.lib()
    ptrs = [C.c_void_p(t.data_ptr()) for t in ins + outs]
    tp = C.c_void_p(tot.data_ptr()) if tot is not None else C.c_void_p(0)
    return getattr(L, fn)(*ptrs, tp, n, C.c_void_p(stream.cuda_stream))


def _run_both(n, dev, torch):
    ins = _tables(n, dev, torch)
    stream = torch.cuda.current_stream()
    res = {}
    for fn in ("nt2h_bs_fwd_adj_f32", "nt2h_bs_fwd_adj_f32_noalloc"):
        outs = [torch.full((n,), float("nan"), dtype=torch.float32, device=dev) for _ in range(6)]
        tot = torch.full((2,), float("nan"), dtype=torch.float64, device=dev)
        print(f"{fn} n={n}", flush=True)
        assert _call(fn, ins, outs, tot, n, stream) == 0
        torch.cuda.synchronize()
        res[fn] = ([o.cpu().numpy() for o in outs], tot.cpu().numpy())
    return res


def _assert_same(res):
    (o1, t1), (o2, t2) = res["nt2h_bs_fwd_adj_f32"], res["nt2h_bs_fwd_adj_f32_noalloc"]
    for a, b in zip(o1, o2):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert np.array_equal(t1.view(np.uint64), t2.view(np.uint64)), (t1, t2)


def test_noalloc_matches_allocating_entry_point_two_calls(dev):
    import torch

    # two calls in a row on the shared scratch, the second with a smaller input
    for n in ((1 << 20) + 3, 4099):
        _assert_same(_run_both(n, dev, torch))


def test_noalloc_at_its_maximum_size_and_beyond(dev):
    import torch

    n = MAX_OPTIONS
    _assert_same(_run_both(n, dev, torch))
    # one option more is refused before anything is launched
    big = [torch.zeros(n + 4, dtype=torch.float32, device=dev) for _ in range(12)]
    tot = torch.zeros(2, dtype=torch.float64, device=dev)
    print("noalloc n=2^24+1 (refused)", flush=True)
    st = _call("nt2h_bs_fwd_adj_f32_noalloc", big[:6], big[6:], tot, n + 1, torch.cuda.current_stream())
    assert st == -7  # NT2H_ERR_TOO_LARGE
    # and a smaller call after the refusal still matches
    _assert_same(_run_both(1000, dev, torch))


def test_noalloc_refuses_host_totals_and_handles_empty(dev):
    import torch

    n = 64
    ins = _tables(n, dev, torch)
    outs = [torch.empty(n, dtype=torch.float32, device=dev) for _ in range(6)]
    host_tot = (C.c_double * 2)()
    L = nt2_amd.lib()
    ptrs = [C.c_void_p(t.data_ptr()) for t in ins + outs]
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    print("noalloc host totals (refused)", flush=True)
    assert L.nt2h_bs_fwd_adj_f32_noalloc(*ptrs, C.cast(host_tot, C.c_void_p), n, s) == -1
    tot = torch.full((2,), 7.0, dtype=torch.float64, device=dev)
    print("noalloc n=0", flush=True)
    assert _call("nt2h_bs_fwd_adj_f32_noalloc", ins, outs, tot, 0, torch.cuda.current_stream()) == 0
    torch.cuda.synchronize()
    assert tot.cpu().tolist() == [0.0, 0.0]
    # no totals: outputs only
    outs2 = [torch.empty(n, dtype=torch.float32, device=dev) for _ in range(6)]
    assert _call("nt2h_bs_fwd_adj_f32", ins, outs2, None, n, torch.cuda.current_stream()) == 0
    assert _call("nt2h_bs_fwd_adj_f32_noalloc", ins, outs, None, n, torch.cuda.current_stream()) == 0
    torch.cuda.synchronize()
    for a, b in zip(outs, outs2):
        assert torch.equal(a.view(torch.int32), b.view(torch.int32))


def test_noalloc_unaligned_tables(dev):
    """Tables one element past a 16-byte boundary (the element-wise kernel,
    whose grid the entry point caps to its scratch): the six streams
    bit-identical to nt2h_bs_fwd_adj_f32's on the same tables, the totals
    (summed over a different number of workgroups) within 1e-12 of the sum
    of |terms| of each, on two calls in a row, the second smaller."""
    import torch

    for n in (100003, 4101):
        d = synth.bs_inputs(n, seed=5, with_w=True)
        ins = [torch.from_numpy(np.concatenate([[0], d[k]]).astype(np.float32)).to(dev)[1:]
               for k in ("S", "K", "T", "r", "sig", "w")]
        assert all(t.data_ptr() % 16 != 0 for t in ins)
        res = {}
        for fn in ("nt2h_bs_fwd_adj_f32", "nt2h_bs_fwd_adj_f32_noalloc"):
            outs = [torch.full((n + 1,), float("nan"), dtype=torch.float32, device=dev)[1:] for _ in range(6)]
            tot = torch.full((2,), float("nan"), dtype=torch.float64, device=dev)
            print(f"{fn} unaligned n={n}", flush=True)
            assert _call(fn, ins, outs, tot, n, torch.cuda.current_stream()) == 0
            torch.cuda.synchronize()
            res[fn] = ([o.cpu().numpy() for o in outs], tot.cpu().numpy())
        (o1, t1), (o2, t2) = res["nt2h_bs_fwd_adj_f32"], res["nt2h_bs_fwd_adj_f32_noalloc"]
        for a, b in zip(o1, o2):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
        for k, j in ((0, 4), (1, 5)):
            mag = np.abs(o1[j].astype(np.float64)).sum()
            assert abs(t1[k] - t2[k]) <= 1e-12 * mag, (k, t1[k], t2[k])


def test_noalloc_recorded_graph_replays(dev):
    import torch

    n = (1 << 18) + 1
    ins = _tables(n, dev, torch)
    ref_outs = [torch.empty(n, dtype=torch.float32, device=dev) for _ in range(6)]
    ref_tot = torch.empty(2, dtype=torch.float64, device=dev)
    assert _call("nt2h_bs_fwd_adj_f32", ins, ref_outs, ref_tot, n, torch.cuda.current_stream()) == 0
    torch.cuda.synchronize()
    outs = [torch.empty(n, dtype=torch.float32, device=dev) for _ in range(6)]
    tot = torch.empty(2, dtype=torch.float64, device=dev)
    g = torch.cuda.CUDAGraph()
    print("noalloc graph capture", flush=True)
    with torch.cuda.graph(g):
        st = _call("nt2h_bs_fwd_adj_f32_noalloc", ins, outs, tot, n, torch.cuda.current_stream())
    assert st == 0
    for rep in range(3):
        for o in outs:
            o.fill_(float("nan"))
        tot.fill_(float("nan"))
        torch.cuda.synchronize()
        print(f"noalloc graph replay {rep}", flush=True)
        g.replay()
        torch.cuda.synchronize()
        for a, b in zip(outs, ref_outs):
            assert torch.equal(a.view(torch.int32), b.view(torch.int32))
        assert torch.equal(tot.view(torch.int64), ref_tot.view(torch.int64))
        # eager work on the same scratch between replays
        small = _tables(333, dev, torch)
        so = [torch.empty(333, dtype=torch.float32, device=dev) for _ in range(6)]
        stot = torch.empty(2, dtype=torch.float64, device=dev)
        assert _call("nt2h_bs_fwd_adj_f32_noalloc", small, so, stot, 333, torch.cuda.current_stream()) == 0
        torch.cuda.synchronize()
