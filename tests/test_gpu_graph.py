"""Entry points recorded in a HIP graph (torch.cuda.CUDAGraph over the C ABI).

A serving loop or a launch-bound caller captures its calls once and replays
them.  The library's kernels reset their own tickets (no memset nodes) and
take every buffer from the caller or from workspaces allocated before the
capture, so replays must reproduce the eager results bit for bit: each test
replays several times, with eager work and synchronisations in between, and
checks every replay.
"""
import numpy as np
import pytest

import nt2_amd
from nt2_amd import synth

pytestmark = pytest.mark.gpu


def _st():
    import torch

    return torch.cuda.current_stream().cuda_stream


def _capture(fn):
    """Record fn on a side stream after running it once there: the library
    keeps its workspaces per (device, stream), and allocation cannot be
    recorded."""
    import torch

    cs = torch.cuda.Stream()
    cs.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(cs):
        fn()
    cs.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=cs):
        fn()
    torch.cuda.synchronize()
    return g


def test_capture_without_workspace_is_refused_cleanly(dev):
    """A call that would have to allocate its workspace while the stream is
    capturing returns NT2H_ERR_CAPTURE before touching the stream; the capture
    itself stays valid."""
    import torch

    x = torch.ones(1 << 16, device=dev)
    out = torch.zeros(1, dtype=torch.float64, device=dev)
    cs = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=cs):
        rc = nt2_amd.lib().nt2h_sum_f32(x.data_ptr(), x.numel(), nt2_amd.SUM_KAHAN, out.data_ptr(), _st())
    torch.cuda.synchronize()
    assert rc == -8 and nt2_amd.STATUS[rc] == "NT2H_ERR_CAPTURE"
    g.replay()  # an empty graph
    torch.cuda.synchronize()


def test_graph_saxpy_and_dot_replays(dev):
    import torch

    n = (1 << 20) + 5
    x_np, y_np = synth.saxpy_inputs(n)
    x = torch.from_numpy(x_np).to(dev)
    y0 = torch.from_numpy(y_np).to(dev)
    y = y0.clone()
    out = torch.zeros(1, dtype=torch.float64, device=dev)

    def step():
        nt2_amd.call("nt2h_saxpy_f32", 2.5, x.data_ptr(), y.data_ptr(), n, _st())
        nt2_amd.call("nt2h_dot_f32", x.data_ptr(), y.data_ptr(), n, nt2_amd.SUM_KAHAN, out.data_ptr(), _st())

    step()
    torch.cuda.synchronize()
    want_y, want_d = y.clone(), out.clone()
    y.copy_(y0)
    g = _capture(step)
    for _ in range(4):
        y.copy_(y0)  # eager work between replays
        torch.cuda.synchronize()
        out.zero_()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(y, want_y)
        assert out.item() == want_d.item()


def test_graph_bs_fwd_adj_with_totals_replays(dev):
    import torch

    n = (1 << 20) + 3
    d = synth.bs_inputs(n, with_w=True)
    ins = [torch.from_numpy(d[k]).to(dev) for k in ("S", "K", "T", "r", "sig", "w")]
    outs = [torch.empty(n, dtype=torch.float32, device=dev) for _ in range(6)]
    tot = torch.zeros(2, dtype=torch.float64, device=dev)

    def step():
        nt2_amd.call("nt2h_bs_fwd_adj_f32", *[t.data_ptr() for t in ins + outs], tot.data_ptr(), n, _st())

    step()
    torch.cuda.synchronize()
    want = [o.clone() for o in outs]
    want_t = tot.clone()
    g = _capture(step)
    for _ in range(4):
        for o in outs:
            o.fill_(-1.0)
        tot.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for o, w in zip(outs, want):
            assert torch.equal(o, w)
        assert torch.equal(tot, want_t)


def test_graph_fold_and_scan_replays(dev):
    import torch

    lo, mid, hi = 1, (1 << 22) + 9, 1  # segmented whole-table paths (several launches)
    x = torch.from_numpy(np.random.default_rng(4).standard_normal(lo * mid * hi).astype(np.float32)).to(dev)
    f = torch.empty(1, dtype=torch.float32, device=dev)
    y = torch.empty_like(x)

    def step():
        nt2_amd.call("nt2h_fold_f32", nt2_amd.FOLD_OPS["norm2"], x.data_ptr(), lo, mid, hi, f.data_ptr(), _st())
        nt2_amd.call("nt2h_scan_f32", nt2_amd.SCAN_OPS["cumsum"], x.data_ptr(), lo, mid, hi, y.data_ptr(), _st())

    step()
    torch.cuda.synchronize()
    wf, wy = f.clone(), y.clone()
    g = _capture(step)
    for _ in range(4):
        f.zero_()
        y.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(f, wf) and torch.equal(y, wy)


def test_graph_lookback_scan_replays(dev):
    """The single-launch look-back scan keeps per-stream tile states tagged
    with a launch epoch that the kernel itself advances: replays, with eager
    calls of the same scan on the same stream between them, give the eager
    bits every time."""
    import torch

    lo, mid, hi = 1, 1 << 18, 2
    x = torch.from_numpy(np.random.default_rng(5).standard_normal(lo * mid * hi).astype(np.float32)).to(dev)
    y = torch.empty_like(x)
    m = torch.empty_like(x)

    def step():
        nt2_amd.call("nt2h_scan_f32", nt2_amd.SCAN_OPS["cumsum"], x.data_ptr(), lo, mid, hi, y.data_ptr(), _st())
        nt2_amd.call("nt2h_scan_f32", nt2_amd.SCAN_OPS["cummax"], x.data_ptr(), lo, mid, hi, m.data_ptr(), _st())

    step()
    torch.cuda.synchronize()
    wy, wm = y.clone(), m.clone()
    cs = torch.cuda.Stream()  # the stream whose state the graph's kernels use
    cs.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(cs):
        step()
    cs.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=cs):
        step()
    torch.cuda.synchronize()
    for _ in range(4):
        y.zero_()
        m.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(y, wy) and torch.equal(m, wm)
        y.zero_()
        torch.cuda.synchronize()
        with torch.cuda.stream(cs):  # eager launches on the same state between replays
            step()
        cs.synchronize()
        assert torch.equal(y, wy) and torch.equal(m, wm)


def test_graph_empty_portfolio_totals(dev):
    """The totals of an empty portfolio are written by a kernel, so the call
    can be recorded: every replay zeroes them again."""
    import torch

    tot = torch.empty(2, dtype=torch.float64, device=dev)

    def step():
        nt2_amd.call("nt2h_bs_fwd_adj_f32", *([None] * 12), tot.data_ptr(), 0, _st())

    g = _capture(step)
    for _ in range(3):
        tot.fill_(float("nan"))
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert tot.cpu().tolist() == [0.0, 0.0]
