"""Tables past 2^31 elements: every index path is 64-bit (NT2's loops run over
std::size_t, core/functions/transform.hpp:55-69).  Checks use inputs whose
results are known exactly, so the whole output can be compared on the GPU."""
import numpy as np
import pytest

import nt2_amd
from nt2_amd import ops

pytestmark = pytest.mark.gpu

N = (1 << 31) + 7  # 8 GiB + 28 bytes of float


@pytest.fixture
def big(dev):
    import torch

    yield
    torch.cuda.empty_cache()


def test_map_and_saxpy_past_2_31(dev, big):
    import torch

    x = torch.arange(N, dtype=torch.float32, device=dev)  # float(i): exact to 2^24, rounded above
    y = ops.map("sqrt", x, nt2_amd.EXACT)  # correctly rounded, as torch.sqrt
    assert torch.equal(y, torch.sqrt(x))
    del y
    ones = torch.ones(N, dtype=torch.float32, device=dev)
    ops.saxpy_(2.0, ones, x)  # x + 2*1: one rounding either way
    assert torch.equal(x[-8:], torch.arange(N - 8, N, dtype=torch.float32, device=dev) + 2.0)
    assert x[0].item() == 2.0 and x[(1 << 31) - 1].item() == float(np.float32((1 << 31) - 1) + np.float32(2))


def test_folds_and_scans_past_2_31(dev, big):
    import torch

    ones = torch.ones(N, dtype=torch.float32, device=dev)
    assert ops.sum(ones).item() == float(N)
    assert ops.dot(ones, ones).item() == float(N)
    g = ops.global_fold("sum", ones).item()
    assert g == float(np.float32(N))
    ones[-1] = 5.0
    ones[17] = float("nan")
    assert ops.global_fold("maximum", ones).item() == 5.0  # after the last NaN, the later max
    ones[-1] = 1.0
    y = ops.scan("cumsum", ones)  # ones with a NaN at 17: NaN from there on
    assert y[16].item() == 17.0 and np.isnan(y[-1].item())
    ones[17] = 1.0
    ops.scan("cumsum", ones, out=y)
    for i in (0, (1 << 24) + 3, (1 << 31) - 1, (1 << 31), N - 1):
        assert y[i].item() == float(np.float32(i + 1)), i  # prefix sums of ones, correctly rounded
    m = ops.scan("cummax", torch.arange(N, dtype=torch.float32, device=dev))
    assert m[-1].item() == float(np.float32(N - 1))


def test_bs_fwd_adj_past_2_31(dev, big):
    import torch

    vals = (10.0, 10.0, 1.0, 0.02, 0.3, 1.0)  # every option the same: outputs all equal
    ins = [torch.full((N,), v, dtype=torch.float32, device=dev) for v in vals]
    outs = ops.bs_fwd_adj(*ins)
    one = ops.bs_fwd_adj(*[t[:64].clone() for t in ins])
    for o, r in zip(outs[:6], one[:6]):
        assert o.min().item() == r[0].item() and o.max().item() == r[0].item()
    dr = float(one[4][0].item())
    assert abs(outs[6][0].item() - N * dr) <= 1e-12 * N * abs(dr)  # compensated totals over 2^31 options
    del ins, outs
