"""Differentiable one-sided Jacobi SVD for PyTorch on the MI355X.

``svd(a)`` takes a CUDA tensor of shape (m, n) or (B, m, n) with m >= n, in
float64 or float32, and returns (U, S, V) with A = U diag(S) V^T: U is
(B, m, n), S is (B, n) in the solver's final column order (not sorted), V is
(B, n, n).  The forward pass is the FP64 (or mixed-precision) Jacobi solve of
libjsvd_amd.so; the backward pass is its hand-written adjoint kernel, in the
dtype of the input.  No reference counterpart: the reference has no Python
binding and no derivative.
"""
from __future__ import annotations

import threading

import torch

from . import BatchPlan, PRECISION_FP64

_plans: dict = {}
_plans_lock = threading.Lock()


def _plan(batch: int, m: int, n: int, epsilon: float, max_sweeps: int, precision: int) -> BatchPlan:
    key = (torch.cuda.current_device(), batch, m, n, float(epsilon), int(max_sweeps), int(precision))
    with _plans_lock:
        plan = _plans.get(key)
        if plan is None:
            plan = BatchPlan([m] * batch, [n] * batch, epsilon=epsilon, max_sweeps=max_sweeps, precision=precision)
            _plans[key] = plan
        return plan


class JacobiSVD(torch.autograd.Function):
    @staticmethod
    def forward(ctx, a, epsilon, max_sweeps, precision):
        if not a.is_cuda:
            raise ValueError("jsvd.autograd.svd needs a CUDA (HIP) tensor")
        if a.dtype not in (torch.float64, torch.float32):
            raise ValueError("jsvd.autograd.svd supports float64 and float32")
        squeeze = a.dim() == 2
        a3 = a.unsqueeze(0) if squeeze else a
        if a3.dim() != 3:
            raise ValueError("input must have shape (m, n) or (B, m, n)")
        b, m, n = a3.shape
        plan = _plan(b, m, n, epsilon, max_sweeps, precision)
        a64 = a3.to(torch.float64).contiguous()
        u = torch.empty((b, m, n), dtype=torch.float64, device=a.device)
        s = torch.empty((b, n), dtype=torch.float64, device=a.device)
        v = torch.empty((b, n, n), dtype=torch.float64, device=a.device)
        plan.execute(a64.view(-1), u.view(-1), s.view(-1), v.view(-1))
        u, s, v = u.to(a.dtype), s.to(a.dtype), v.to(a.dtype)
        ctx.plan = plan
        ctx.squeeze = squeeze
        ctx.save_for_backward(u, s, v)
        if squeeze:
            return u[0], s[0], v[0]
        return u, s, v

    @staticmethod
    def backward(ctx, gu, gs, gv):
        u, s, v = ctx.saved_tensors
        prep = lambda g, like: (None if g is None else
                                g.reshape(like.shape).to(like.dtype).contiguous())
        gu, gs, gv = prep(gu, u), prep(gs, s), prep(gv, v)
        ga = torch.empty_like(u)
        ctx.plan.backward(u.view(-1), s.view(-1), v.view(-1),
                          None if gu is None else gu.view(-1), None if gs is None else gs.view(-1),
                          None if gv is None else gv.view(-1), ga.view(-1))
        if ctx.squeeze:
            ga = ga[0]
        return ga, None, None, None


def svd(a: torch.Tensor, epsilon: float = 1e-15, max_sweeps: int = 30, precision: int = PRECISION_FP64):
    """Differentiable SVD (U, S, V) of a (m, n) or (B, m, n) CUDA tensor."""
    return JacobiSVD.apply(a, epsilon, max_sweeps, precision)
