"""torch.autograd binding of the derivative kernels.

log f(u) of a path state, with the state's triangles fixed (the function of
the reference's chad library, evaluate_path_bidir_{c}_{l}_static,
src/chad.cpp:183-238), as a differentiable function of its primary-sample
vector u. The forward pass evaluates log f and the jet kernels' gradient and
Hessian (drmlt_eval_logf and drmlt_grad_hess of libdrmlt_hip.so); backward
returns the gradient, and the gradient is itself differentiable with the
Hessian as its Jacobian, so double backward (Hessian-vector products) works.
Derivatives beyond the second are not provided.

The tensors are CPU float64: the C ABI copies host buffers in and out, and
all arithmetic happens in the HIP kernels.
"""
from __future__ import annotations

import numpy as np
import torch

from . import capi


def _dims(states):
    return 2 * np.maximum(states["c"] + states["l"] - 1, 2)


def _with_u(states, u):
    recs = np.array(states, capi.PATH_REC, copy=True)
    un = u.detach().cpu().numpy()
    recs["u"][:, :un.shape[1]] = un.astype(np.float32)
    return recs


class _GradLogF(torch.autograd.Function):
    """u -> the gradient computed in the forward pass of LogF; its Jacobian is
    the Hessian computed there."""

    @staticmethod
    def forward(ctx, u, grad, hess):
        ctx.save_for_backward(hess)
        return grad.clone()

    @staticmethod
    def backward(ctx, gg):
        (hess,) = ctx.saved_tensors
        return torch.bmm(hess, gg.unsqueeze(-1)).squeeze(-1), None, None


class LogF(torch.autograd.Function):
    """log f(u) for n states of one dimension D: u is (n, D), the result (n,)."""

    @staticmethod
    def forward(ctx, u, context, states, fp64):
        recs = _with_u(states, u)
        D = int(u.shape[1])
        if not (_dims(recs) == D).all():
            raise ValueError("all states of one call must have the primary-sample dimension of u")
        lf, ok = context.eval_logf(recs, fp64=fp64)
        g, h, okd = context.grad_hess(recs, fp64=fp64)
        bad = (ok == 0) | (okd == 0)
        lf = np.where(bad, np.nan, lf)
        grad = torch.from_numpy(np.where(bad[:, None], np.nan, g[:, :D]))
        hess = torch.from_numpy(np.where(bad[:, None, None], np.nan, h[:, :D, :D]))
        ctx.save_for_backward(u, grad, hess)
        return torch.from_numpy(lf)

    @staticmethod
    def backward(ctx, gout):
        u, grad, hess = ctx.saved_tensors
        g = _GradLogF.apply(u, grad, hess)
        return gout.unsqueeze(-1) * g, None, None, None


def log_f(u: torch.Tensor, context: "capi.Context", states, fp64: bool = True) -> torch.Tensor:
    """Differentiable log f at the states with primary samples u (n, D)."""
    return LogF.apply(u, context, states, fp64)
