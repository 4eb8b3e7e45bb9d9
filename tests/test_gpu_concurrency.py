"""Calls on distinct streams from distinct host threads: the library keeps one
reduction workspace and one scratch buffer per (device, stream), so
concurrent folds must give the same bits as sequential ones."""
import threading

import numpy as np
import pytest

from nt2_amd import ops, synth

pytestmark = pytest.mark.gpu


def test_threads_on_separate_streams_match_sequential(dev):
    import torch

    n = (1 << 20) + 13
    d = synth.bs_inputs(n, with_w=True)
    g = [torch.from_numpy(d[k]).to(dev) for k in ("S", "K", "T", "r", "sig", "w")]
    x = torch.from_numpy(np.random.default_rng(3).standard_normal(1 << 22).astype(np.float32)).to(dev)
    X = x.reshape(1024, 4096)  # NT2 extent 4096 x 1024
    torch.cuda.synchronize()

    def run_all():
        out = ops.bs_fwd_adj(*g)
        return [out[6].cpu().numpy(), ops.sum(x).cpu().numpy(), ops.fold("maximum", X, 2).cpu().numpy(),
                ops.fold("sum", X, 1).cpu().numpy(), ops.scan("cumsum", x).cpu().numpy()[-5:],
                out[0].cpu().numpy()[:1000]]

    want = run_all()
    results, errors = {}, []

    def worker(k):
        try:
            s = torch.cuda.Stream()  # default priority
            with torch.cuda.stream(s):
                got = None
                for _ in range(5):
                    got = run_all()
                    for a, b in zip(got, want):
                        if not np.array_equal(a.view(np.uint8), b.view(np.uint8)):
                            raise AssertionError(f"thread {k}: results differ from the sequential run")
                s.synchronize()
            results[k] = True
        except Exception as e:  # surfaced in the main thread
            errors.append(e)

    threads = [threading.Thread(target=worker, args=(k,)) for k in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not errors, errors[0]
    assert len(results) == 4
