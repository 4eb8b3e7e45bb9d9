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
