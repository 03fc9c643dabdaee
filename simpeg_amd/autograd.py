"""torch.autograd bridge: dpred(m) with Jtvec as its backward.

``dpred(sim, m)`` returns predicted data as a torch tensor whose gradient with respect to the
model ``m`` is ``sim.Jtvec(m, grad)`` -- the hand-written adjoint (batched adjoint solve +
edge-gradient contraction kernel).  With ``recompute=True`` the forward keeps no source fields:
the backward solves the forward problem again (activation checkpointing for large surveys).
"""

from __future__ import annotations

import numpy as np
import torch


class DpredFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, m, sim, recompute):
        m_np = m.detach().to("cpu", torch.float64).numpy().copy()
        f = sim.fields(m_np)
        d = sim.dpred(m_np, f=f)
        ctx.sim = sim
        ctx.m_np = m_np
        ctx.fields = None if recompute else f
        return torch.as_tensor(d, dtype=m.dtype, device=m.device)

    @staticmethod
    def backward(ctx, grad_d):
        sim = ctx.sim
        f = ctx.fields if ctx.fields is not None else sim.fields(ctx.m_np)
        g = sim.Jtvec(ctx.m_np, grad_d.detach().to("cpu", torch.float64).numpy(), f=f)
        return torch.as_tensor(g, dtype=grad_d.dtype, device=grad_d.device), None, None


def dpred(sim, m, recompute=False):
    """Differentiable predicted data of ``sim`` at model ``m`` (torch tensor)."""
    if not isinstance(m, torch.Tensor):
        m = torch.as_tensor(np.asarray(m, dtype=float))
    return DpredFunction.apply(m, sim, recompute)


def half_weighted_misfit(sim, m, dobs, weights, recompute=False):
    """0.5 * ||W (dpred(m) - dobs)||^2 with W = diag(weights); differentiable in m."""
    d = dpred(sim, m, recompute=recompute)
    w = torch.as_tensor(np.asarray(weights, dtype=float), dtype=d.dtype, device=d.device)
    r = w * (d - torch.as_tensor(np.asarray(dobs, dtype=float), dtype=d.dtype, device=d.device))
    return 0.5 * (r * r).sum()
