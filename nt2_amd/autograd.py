"""torch.autograd.Function wrappers over the hand-written adjoint kernels.

They exist to check the adjoints (gradcheck-style against finite
differences) through the same C ABI a C++ caller uses: the backward of
each Function is one launch of the fused forward+adjoint kernel with the
incoming cotangent, recomputing the forward in registers.
"""
from __future__ import annotations

import torch

from . import FUSED, ops


class BlackScholesCall(torch.autograd.Function):
    """price = call(S, K, T, r, sig); backward = nt2h_bs_fwd_adj with w = grad_output."""

    @staticmethod
    def forward(ctx, S, K, T, r, sig):
        call, _ = ops.bs_fwd(S, K, T, r, sig, precision=FUSED)
        ctx.save_for_backward(S, K, T, r, sig)
        return call

    @staticmethod
    def backward(ctx, g):
        S, K, T, r, sig = ctx.saved_tensors
        outs = ops.bs_fwd_adj(S, K, T, r, sig, g.contiguous(), totals=False)
        return outs[1], outs[2], outs[3], outs[4], outs[5]


class ColumnFold(torch.autograd.Function):
    """z_j = sum_i exp(-X_ij^2/(2 s_j^2)) cos(k_j X_ij + p_j) over a (cols, rows) tensor
    (= column-major rows x cols table); gradients for s, k, p (X is data)."""

    @staticmethod
    def forward(ctx, X, s, k, p):
        z = ops.colfold_gc(X, s, k, p, grad=False)
        ctx.save_for_backward(X, s, k, p)
        return z

    @staticmethod
    def backward(ctx, gz):
        X, s, k, p = ctx.saved_tensors
        _, ds, dk, dp = ops.colfold_gc(X, s, k, p, gz=gz.contiguous())
        return None, ds, dk, dp


def bs_call(S, K, T, r, sig):
    return BlackScholesCall.apply(S, K, T, r, sig)


def colfold(X, s, k, p):
    return ColumnFold.apply(X, s, k, p)
