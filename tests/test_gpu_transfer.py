"""Staged host <-> device copies and the host-table Black-Scholes pipeline."""
import ctypes as C

import numpy as np
import pytest

import nt2_amd
from nt2_amd import ops, synth

pytestmark = pytest.mark.gpu


def _stream():
    import torch

    return torch.cuda.current_stream().cuda_stream


@pytest.mark.parametrize("nbytes", [1, 4097, (8 << 20) + 5, (20 << 20) + 3])
def test_upload_download_round_trip(dev, nbytes):
    import torch

    host = np.random.default_rng(nbytes).integers(0, 256, nbytes, dtype=np.uint8)
    d = torch.empty(nbytes + 1, dtype=torch.uint8, device=dev)
    dst = d.data_ptr() + 1  # an unaligned device address
    nt2_amd.call("nt2h_upload", dst, host.ctypes.data, nbytes, _stream())
    assert np.array_equal(d[1:].cpu().numpy(), host)
    back = np.zeros(nbytes, np.uint8)
    nt2_amd.call("nt2h_download", back.ctypes.data, dst, nbytes, _stream())
    assert np.array_equal(back, host)


def _host_adj(d, keys, n, block):
    outs = [np.empty(n, np.float32) for _ in range(6)]
    tot = np.zeros(2)
    nt2_amd.call("nt2h_bs_fwd_adj_f32_host", *[d[k].ctypes.data for k in keys], *[o.ctypes.data for o in outs],
                 tot.ctypes.data, n, block)
    return outs, tot


@pytest.mark.parametrize("n,block", [(3 * (1 << 20) + 5, 1 << 20), (100003, 1 << 20), (70001, 9999), (5, 0)])
def test_bs_host_pipeline_matches_device_kernel(dev, n, block):
    import torch

    d = synth.bs_inputs(n, with_w=True)
    keys = ("S", "K", "T", "r", "sig", "w")
    outs, tot = _host_adj(d, keys, n, block)
    g = [torch.from_numpy(d[k]).to(dev) for k in keys]
    want = ops.bs_fwd_adj(*g)
    for a, b in zip(outs, want[:6]):
        assert np.array_equal(a.view(np.uint32), b.cpu().numpy().view(np.uint32))
    wt = want[6].cpu().numpy()
    single_block = block <= 0 or block >= n
    for i, j in ((0, 4), (1, 5)):
        if single_block:  # one launch over all options: the same combine as the device call
            assert tot[i] == wt[i]
        else:  # block totals summed in block order
            assert abs(tot[i] - wt[i]) <= 1e-13 * np.abs(outs[j].astype(np.float64)).sum()


def test_bs_host_pipeline_empty(dev):
    tot = np.full(2, 7.0)
    z = np.zeros(1, np.float32)
    nt2_amd.call("nt2h_bs_fwd_adj_f32_host", *[z.ctypes.data] * 12, tot.ctypes.data, 0, 0)
    assert tot[0] == 0 and tot[1] == 0


@pytest.mark.parametrize("nbytes", [0, 1, 4099, (3 << 20) + 7])
def test_pinned_and_device_buffers_with_memcpy(dev, nbytes):
    """nt2h_host_alloc (pinned) -> nt2h_memcpy H2D into nt2h_malloc memory ->
    D2D -> D2H and H2H into more pinned memory: the bytes arrive unchanged for
    every copy kind (and DEFAULT), zero-byte copies included; every buffer is
    released through its own free. Sizes past 1 MiB take nt2h_malloc's offset
    (coloured) allocations, whose free must find the base again."""
    L = nt2_amd.lib()
    s = C.c_void_p(_stream())
    if nbytes == 0:  # a zero-byte request succeeds with a null pointer
        for alloc in (L.nt2h_malloc, L.nt2h_host_alloc):
            z = C.c_void_p(1)
            assert alloc(C.byref(z), 0) == 0 and not z.value
    hp = [C.c_void_p() for _ in range(3)]
    dp = [C.c_void_p() for _ in range(2)]
    for p in hp:
        assert L.nt2h_host_alloc(C.byref(p), max(nbytes, 1)) == 0 and p.value
    for p in dp:
        assert L.nt2h_malloc(C.byref(p), max(nbytes, 1)) == 0 and p.value
    try:
        src = np.random.default_rng(nbytes).integers(0, 256, nbytes, dtype=np.uint8)
        C.memmove(hp[0].value, src.ctypes.data, nbytes)
        assert L.nt2h_memcpy(dp[0], hp[0], nbytes, 1, s) == 0  # H2D
        assert L.nt2h_memcpy(dp[1], dp[0], nbytes, 3, s) == 0  # D2D
        assert L.nt2h_memcpy(hp[1], dp[1], nbytes, 2, s) == 0  # D2H
        assert L.nt2h_memcpy(hp[2], hp[1], nbytes, 0, s) == 0  # H2H
        assert L.nt2h_memcpy(hp[1], dp[0], nbytes, 4, s) == 0  # DEFAULT (inferred)
        assert L.nt2h_sync(s) == 0
        for p in hp[1:]:
            got = np.frombuffer((C.c_uint8 * nbytes).from_address(p.value), np.uint8) if nbytes else np.zeros(0, np.uint8)
            assert np.array_equal(got, src)
    finally:
        for p in hp:
            assert L.nt2h_host_free(p) == 0
        for p in dp:
            assert L.nt2h_free(p) == 0
