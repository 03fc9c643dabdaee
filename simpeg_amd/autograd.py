"""torch.autograd bridge: dpred(m) with Jtvec as its backward.

``dpred(sim, m)`` returns predicted data as a torch tensor whose gradient with respect to the
model ``m`` is ``sim.Jtvec(m, grad)`` -- the hand-written adjoint (batched adjoint solve +
edge-gradient contraction kernel).  With ``recompute=True`` the forward keeps no source fields:
the backward solves the forward problem again (activation checkpointing for large surveys).

A model in a GPU tensor stays on the GPU when the simulation's map evaluates on device tensors
(Identity, Exp and Reciprocal maps, no miniaturisation): the data, the incoming gradient and
the model gradient are device tensors and nothing is copied to the host.  Other maps go through
the host API (numpy) and copy the model, data and gradient.
"""

from __future__ import annotations

import numpy as np
import torch


class DpredFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, m, sim, recompute):
        ctx.sim = sim
        if m.is_cuda and sim._device_map() is not None:
            d, state = sim.dpred_device(m)
            ctx.device = True
            ctx.m = m.detach()
            ctx.state = None if recompute else state
            return d.to(m.dtype)
        ctx.device = False
        m_np = m.detach().to("cpu", torch.float64).numpy().copy()
        f = sim.fields(m_np)
        d = sim.dpred(m_np, f=f)
        ctx.m_np = m_np
        ctx.fields = None if recompute else f
        return torch.as_tensor(d, dtype=m.dtype, device=m.device)

    @staticmethod
    def backward(ctx, grad_d):
        sim = ctx.sim
        if ctx.device:
            g = sim.jtvec_device(ctx.m, grad_d, state=ctx.state)
            return g.to(grad_d.dtype), None, None
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
    w = torch.as_tensor(weights, dtype=d.dtype).to(d.device)
    r = w * (d - torch.as_tensor(dobs, dtype=d.dtype).to(d.device))
    return 0.5 * (r * r).sum()
