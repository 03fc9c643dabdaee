"""Gauss-Newton data fitting with the GPU nodal DC simulation.

Objective phi(m) = 1/2 ||Wd (dpred(m) - dobs)||^2 + beta/2 ||Wm (m - mref)||^2.  Each
Gauss-Newton step solves (J^T Wd^T Wd J + beta Wm^T Wm) dm = -grad with a fixed number of
conjugate-gradient steps (the inexact Gauss-Newton of simpeg/optimization.py:1345-1416, without
its BFGS preconditioner), every Hessian product being one batched Jvec and one batched Jtvec,
followed by an Armijo backtracking line search.  With ``recompute=True`` the source fields are not
kept between Hessian products (checkpointing): each product solves the forward problem once and
shares those fields between its Jvec and its Jtvec, then drops them, so the fields never
outlive one product; the gradient reuses the fields of the line search's accepted model.  The
peak device memory is that of one product (fields, adjoint block and solver workspace) either
way; recomputing costs one extra batched solve per product.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass
class GNHistory:
    phi_d: list = field(default_factory=list)
    phi_m: list = field(default_factory=list)
    phi: list = field(default_factory=list)
    step: list = field(default_factory=list)
    cg_rel_res: list = field(default_factory=list)


def gauss_newton(sim, dobs, wd, m0, beta, mref=None, wm=None, n_iter=3, cg_iter=10,
                 recompute=False, ls_max=8, armijo=1e-4, verbose=False):
    """Run ``n_iter`` Gauss-Newton iterations; returns (m, GNHistory).

    wd, wm: diagonal weights (arrays); wm defaults to ones, mref to m0.
    """
    dobs = np.asarray(dobs, dtype=float)
    wd = np.asarray(wd, dtype=float)
    m = np.asarray(m0, dtype=float).copy()
    mref = m.copy() if mref is None else np.asarray(mref, dtype=float)
    wm = np.ones_like(m) if wm is None else np.asarray(wm, dtype=float)
    wm2 = wm * wm
    wd2 = wd * wd
    hist = GNHistory()

    def objective(mm, f=None):
        if f is None:
            f = sim.fields(mm)
        r = sim.dpred(mm, f=f) - dobs
        pd = 0.5 * float(np.sum((wd * r) ** 2))
        pm = 0.5 * beta * float(np.sum((wm * (mm - mref)) ** 2))
        return pd, pm, r, f

    pd, pm, r, f = objective(m)
    hist.phi_d.append(pd)
    hist.phi_m.append(pm)
    hist.phi.append(pd + pm)
    for it in range(n_iter):
        # f: the fields at m from the last objective evaluation
        g = sim.Jtvec(m, wd2 * r, f=f) + beta * wm2 * (m - mref)
        if recompute:
            f = None

        def H(v):
            fh = sim.fields(m) if recompute else f
            jv = sim.Jvec(m, v, f=fh)
            return sim.Jtvec(m, wd2 * jv, f=fh) + beta * wm2 * v

        # conjugate gradients on H dm = -g, fixed number of steps
        dm = np.zeros_like(m)
        res = -g.copy()
        p = res.copy()
        rr = res @ res
        rr0 = rr
        for _ in range(cg_iter):
            hp = H(p)
            al = rr / (p @ hp)
            dm += al * p
            res -= al * hp
            rr_new = res @ res
            p = res + (rr_new / rr) * p
            rr = rr_new
        hist.cg_rel_res.append(float(np.sqrt(rr / rr0)))
        # Armijo backtracking
        t = 1.0
        phi0 = pd + pm
        slope = float(g @ dm)
        for _ in range(ls_max):
            pd_t, pm_t, r_t, f_t = objective(m + t * dm)
            if pd_t + pm_t <= phi0 + armijo * t * slope:
                break
            t *= 0.5
        m = m + t * dm
        pd, pm, r, f = pd_t, pm_t, r_t, f_t
        hist.phi_d.append(pd)
        hist.phi_m.append(pm)
        hist.phi.append(pd + pm)
        hist.step.append(t)
        if verbose:
            print(f"GN {it + 1}: phi_d {pd:.6e} phi_m {pm:.6e} step {t} cg rel res "
                  f"{hist.cg_rel_res[-1]:.2e}", flush=True)
    return m, hist
