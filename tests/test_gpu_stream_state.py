"""Per-stream state of the library: the look-back scans' bounded-wait error
path (forced from the host so that every kernel still ends), scratch growth
under a recorded graph, and releasing the state of streams that come and go."""
import ctypes as C

import numpy as np
import pytest

import nt2_amd

pytestmark = pytest.mark.gpu


def _lib():
    return nt2_amd.lib()


def _scan(x, y, stream):
    # one run of 2^16 floats: the single-launch look-back path
    return _lib().nt2h_scan_f32(nt2_amd.SCAN_OPS["cumsum"], C.c_void_p(x.data_ptr()), 1, x.numel(), 1,
                                C.c_void_p(y.data_ptr()), C.c_void_p(stream))


def test_lookback_wait_that_gives_up_is_reported_and_cleared(dev, ref):
    import torch

    n = 1 << 16
    xh = np.random.default_rng(5).uniform(-1, 1, n).astype(np.float32)
    x = torch.from_numpy(xh).to(dev)
    y = torch.empty_like(x)
    st = torch.cuda.current_stream().cuda_stream
    L = _lib()
    assert _scan(x, y, st) == 0
    assert L.nt2h_sync(C.c_void_p(st)) == 0
    good = y.cpu().numpy().copy()
    try:
        # every waiting tile reads its predecessors as "not yet" and gives up
        # after two polls: poisoned states, the kernel ends, the error is kept
        assert L.nt2h_test_set_lookback_wait(2, 1) == 0
        print("look-back scan with waits that give up", flush=True)
        assert _scan(x, y, st) == 0
        rc = L.nt2h_sync(C.c_void_p(st))
        assert rc == -9 and nt2_amd.STATUS[rc] == "NT2H_ERR_WAIT"
        assert L.nt2h_sync(C.c_void_p(st)) == 0  # reported once, then cleared
    finally:
        L.nt2h_test_set_lookback_wait(0, 0)
    # the state is consistent again: the next call is right and reports nothing
    y.zero_()
    assert _scan(x, y, st) == 0
    assert L.nt2h_sync(C.c_void_p(st)) == 0
    assert np.array_equal(y.cpu().numpy().view(np.uint32), good.view(np.uint32))
    want = ref.scan_lmh("cumsum", xh, 1, n, 1).astype(np.float64)
    assert np.max(np.abs(good.astype(np.float64) - want)) <= 1e-4 * np.max(np.abs(want)) + 1e-6


def test_graph_replay_after_scratch_grew(dev):
    """Capture a segmented fold, grow the stream's scratch with a larger eager
    fold, then replay: the graph keeps its buffer (growth retires, never
    frees), so every replay matches the eager result bit for bit."""
    import torch

    cs = torch.cuda.Stream()
    small = torch.from_numpy(np.random.default_rng(6).standard_normal(1 << 16).astype(np.float32)).to(dev)
    out = torch.zeros(1, dtype=torch.float32, device=dev)
    big = torch.from_numpy(np.random.default_rng(7).standard_normal(4097 * 2048).astype(np.float32)).to(dev)
    bout = torch.zeros(4097, dtype=torch.float32, device=dev)

    def fold_small():
        nt2_amd.call("nt2h_fold_f32", nt2_amd.FOLD_OPS["sum"], C.c_void_p(small.data_ptr()), 1, small.numel(), 1,
                     C.c_void_p(out.data_ptr()), C.c_void_p(cs.cuda_stream))

    cs.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(cs):
        fold_small()
    cs.synchronize()
    want = out.clone()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=cs):
        fold_small()
    torch.cuda.synchronize()
    # 4097 outputs of 2048 rows: the scalar column path, 8 MiB of segment summaries
    with torch.cuda.stream(cs):
        print("eager fold that grows the capture stream's scratch", flush=True)
        nt2_amd.call("nt2h_fold_f32", nt2_amd.FOLD_OPS["sum"], C.c_void_p(big.data_ptr()), 4097, 2048, 1,
                     C.c_void_p(bout.data_ptr()), C.c_void_p(cs.cuda_stream))
    cs.synchronize()
    bwant = big.view(2048, 4097).double().sum(0)
    assert torch.allclose(bout.double(), bwant, rtol=1e-5, atol=1e-3)
    for rep in range(3):
        out.fill_(float("nan"))
        torch.cuda.synchronize()
        print(f"replay {rep}", flush=True)
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(out.view(torch.int32), want.view(torch.int32))
    del g
    torch.cuda.synchronize()
    assert _lib().nt2h_release_stream(C.c_void_p(cs.cuda_stream)) == 0


def test_stream_churn_releases_state(dev):
    """Streams that come and go, each used by every kind of call that keeps
    per-stream state: with nt2h_release_stream after each, device memory
    does not grow, and a handle used again starts from fresh state (same
    results bit for bit)."""
    import torch

    from nt2_amd import synth

    n = 1 << 18
    d = synth.bs_inputs(n, with_w=True)
    ins = [torch.from_numpy(d[k]).to(dev) for k in ("S", "K", "T", "r", "sig", "w")]
    outs = [torch.empty(n, device=dev) for _ in range(6)]
    tot = torch.empty(2, dtype=torch.float64, device=dev)
    x = torch.from_numpy(np.random.default_rng(8).standard_normal(1 << 16).astype(np.float32)).to(dev)
    y = torch.empty_like(x)
    s1 = torch.empty(1, dtype=torch.float64, device=dev)
    col = torch.empty(8, dtype=torch.float32, device=dev)
    L = _lib()
    results = []

    def use(handle):
        hs = C.c_void_p(handle)
        nt2_amd.call("nt2h_bs_fwd_adj_f32", *[C.c_void_p(t.data_ptr()) for t in ins + outs],
                     C.c_void_p(tot.data_ptr()), n, hs)
        nt2_amd.call("nt2h_sum_f32", C.c_void_p(x.data_ptr()), x.numel(), nt2_amd.SUM_KAHAN,
                     C.c_void_p(s1.data_ptr()), hs)
        nt2_amd.call("nt2h_fold_f32", nt2_amd.FOLD_OPS["maximum"], C.c_void_p(x.data_ptr()), 8, 8192, 1,
                     C.c_void_p(col.data_ptr()), hs)
        assert _scan(x, y, handle) == 0
        assert L.nt2h_sync(hs) == 0
        results.append((tot.cpu().numpy().copy(), s1.cpu().numpy().copy(), y.cpu().numpy().copy()))

    torch.cuda.synchronize()

    def held():
        b = C.c_size_t()
        assert L.nt2h_state_bytes(C.byref(b)) == 0
        return b.value

    # torch hands out streams from a per-device pool, so handles come back
    # (earlier tests may have left state on them: released first); each use
    # after a release must start from fresh state
    for it in range(48):
        stream = torch.cuda.Stream()
        assert L.nt2h_release_stream(C.c_void_p(stream.cuda_stream)) == 0
        b0 = held()
        use(stream.cuda_stream)
        b1 = held()
        assert L.nt2h_release_stream(C.c_void_p(stream.cuda_stream)) == 0
        b2 = held()
        # workspace (~4.3 MiB), scratch and look-back state (1 MiB each):
        # allocated by the calls, returned in full by the release
        assert b1 - b0 >= 6 << 20, (b0, b1)
        assert b2 == b0, (b0, b2)
        del stream
        torch.cuda.synchronize()
    for r in results[1:]:
        for a, b in zip(r, results[0]):
            assert np.array_equal(a.view(np.uint8), b.view(np.uint8))
