"""GPU checks of the C++ front end (include/nt2hip/table.hpp) and of the
torch.autograd.Function wrappers over the adjoint kernels."""
import os
import subprocess

import numpy as np
import pytest

import nt2_amd
from nt2_amd import autograd, synth

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_cpp_table_frontend(dev):
    exe = os.path.join(ROOT, "tests", "cpp", "test_table")
    assert os.path.exists(exe), "build tests/cpp (make -C tests/cpp) before the GPU run"
    print("running", exe, flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    print(r.stderr[-2000:])
    assert r.returncode == 0
    assert "FAIL" not in r.stdout
    assert r.stdout.count("PASS") >= 40


def test_autograd_bs_call_vs_central_differences(dev, ref):
    import torch

    n = 2048
    d = synth.bs_inputs(n, with_w=True)
    keys = ("S", "K", "T", "r", "sig")
    ts = [torch.from_numpy(d[k]).to(dev).requires_grad_(True) for k in keys]
    w = torch.from_numpy(d["w"]).to(dev)
    price = autograd.bs_call(*ts)
    (price * w).sum().backward()
    d64 = {k: d[k].astype(np.float64) for k in keys}
    for j, k in enumerate(keys):
        h = 1e-3 * np.abs(d64[k])
        up = dict(d64); up[k] = d64[k] + h
        dn = dict(d64); dn[k] = d64[k] - h
        cu, _ = ref.bs_fwd(*[up[q] for q in keys], dtype=np.float64)
        cd, _ = ref.bs_fwd(*[dn[q] for q in keys], dtype=np.float64)
        fd = d["w"].astype(np.float64) * (cu - cd) / (2 * h)
        got = ts[j].grad.cpu().numpy().astype(np.float64)
        scale = np.abs(d["w"]) * np.maximum(np.abs(fd / np.maximum(np.abs(d["w"]), 1e-30)), 1e-3 * d64["S"])
        assert np.max(np.abs(got - fd) / scale) < 2e-3, k


def test_autograd_colfold_vs_central_differences(dev):
    import torch

    rows, cols = 4096, 32
    X_np, s_np, k_np, p_np = synth.colfold_inputs(rows, cols)
    X = torch.from_numpy(X_np).to(dev)
    s, k, p = (torch.from_numpy(a).to(dev).requires_grad_(True) for a in (s_np, k_np, p_np))
    gz = torch.from_numpy(synth.uniform(0, 99, cols, -1.0, 1.0)).to(dev)
    z = autograd.colfold(X, s, k, p)
    (z * gz).sum().backward()

    def zf(sv, kv, pv):  # FP64 forward on the GPU with torch's own math
        Xd = X.double()
        return (torch.exp(-Xd * Xd / (2 * sv[:, None] ** 2)) * torch.cos(kv[:, None] * Xd + pv[:, None])).sum(1)

    s64, k64, p64 = s.detach().double(), k.detach().double(), p.detach().double()
    for name, grad, base in (("s", s.grad, s64), ("k", k.grad, k64), ("p", p.grad, p64)):
        h = 1e-4 * torch.clamp(base.abs(), min=1.0)
        args_up = [s64, k64, p64]
        args_dn = [s64, k64, p64]
        i = "skp".index(name)
        args_up[i] = base + h
        args_dn[i] = base - h
        fd = gz.double() * (zf(*args_up) - zf(*args_dn)) / (2 * h)
        e = torch.exp(-X.double() ** 2 / (2 * s64[:, None] ** 2))
        mag = (e * (1 + X.double().abs() + X.double() ** 2)).sum(1) * gz.double().abs() / torch.clamp(s64, max=1.0) ** 3
        err = ((grad.double() - fd).abs() / mag).max().item()
        assert err < 1e-5, (name, err)
