"""Autotuner for the multigrid / launch parameters of the nodal DC solver.

For a given problem (engine with sigma set, right-hand-side block) it times complete batched
solves for a small grid of candidates -- smoothing schedule (nu_pre, nu_post), Chebyshev
interval ratio, W-cycle level, interpolation, poll cadence, columns per work-item -- and keeps
the fastest that converges.  Results are
cached in a JSON file keyed by mesh shape, column count (rounded up to a power of two), boundary
condition and precision, so a later run with the same shape reuses the choice without timing.
"""

from __future__ import annotations

import json
import os
import time

import torch

DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "simpeg_amd", "autotune.json")

# Every candidate is a symmetric cycle (nu_pre == nu_post), as PCG needs a symmetric
# preconditioner.  Fields not named keep the engine's current values.
_BASE = dict(nu_pre=2, nu_post=2, cheb_ratio=8.0, rhs_per_thread=0, xfer_rhs_per_thread=0,
             wcycle_level=-1, wcycle_smooth=1, interp=1, vcycle_bf16=1, check_every=1)
CANDIDATES = [
    dict(_BASE),
    dict(_BASE, wcycle_level=0),
    dict(_BASE, wcycle_level=2),
    dict(_BASE, wcycle_smooth=2),
    dict(_BASE, interp=0),
    dict(_BASE, cheb_ratio=4.0),
    dict(_BASE, nu_pre=3, nu_post=3),
    dict(_BASE, check_every=2),
    dict(_BASE, check_every=4),
    dict(_BASE, rhs_per_thread=8, xfer_rhs_per_thread=8),
    dict(_BASE, vcycle_bf16=2),
    # few columns: the fused coarse tail runs one work-group per column, so it can idle most
    # of the device; the multi-launch coarse levels spread the same work over nodes
    dict(_BASE, fused_tail=0),
    dict(_BASE, fused_tail=0, wcycle_level=0),
]


def _key(engine, ncol, mode):
    # column counts share an entry up to the next power of two (the kernels' columns per
    # work-item and the relative cost of the coarse levels change at that granularity)
    bucket = 1
    while bucket < ncol:
        bucket *= 2
    return f"{engine.shape_cells}|ncol<={bucket}|bc={engine.bc}|mode={mode}|levels={engine.levels}"


def load_cache(path=DEFAULT_CACHE):
    try:
        with open(path) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return {}


def save_cache(cache, path=DEFAULT_CACHE):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as fh:
        json.dump(cache, fh, indent=1, sort_keys=True)
    os.replace(tmp, path)


def tune(engine, q, ncol=None, mode="mixed", tol=1e-10, candidates=None, repeats=1,
         cache_path=DEFAULT_CACHE, use_cache=True):
    """Pick and apply the fastest parameter set for solving ``q`` on ``engine``.

    Returns (params dict, timings list).  ``engine.set_sigma`` must have been called.
    """
    ncol = q.shape[0] if ncol is None else ncol
    key = _key(engine, ncol, mode)
    cache = load_cache(cache_path) if use_cache else {}
    # the candidate's fields override the engine's current parameters; the rest (e.g. graph,
    # vcycle storage) stay as the caller set them, also when the choice comes from the cache
    base = {f: getattr(engine.params, f) for f, _ in engine.params._fields_}
    if use_cache and key in cache:
        best = {**base, **cache[key]["params"]}
        engine.set_params(**best)
        return best, cache[key].get("timings", [])
    cands = candidates or CANDIDATES
    timings = []
    best, best_cand, best_t = None, None, float("inf")
    u = engine.block(ncol, zero=False)
    for cand in cands:
        p = {**base, **cand}
        engine.set_params(**p)
        try:
            engine.solve(q, ncol, mode=mode, tol=tol, u=u)  # warm-up (workspace, caches)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(repeats):
                _, res = engine.solve(q, ncol, mode=mode, tol=tol, u=u)
            torch.cuda.synchronize()
            dt = (time.perf_counter() - t0) / repeats
        except RuntimeError:
            continue
        timings.append(dict(params=cand, seconds=dt, iters=res.iters, fp64_iters=res.fp64_iters))
        if dt < best_t:
            best, best_cand, best_t = p, cand, dt
    if best is None:
        engine.set_params(**base)
        raise RuntimeError("no autotuner candidate converged")
    engine.set_params(**best)
    if use_cache:
        cache[key] = dict(params=best_cand, timings=timings)
        save_cache(cache, cache_path)
    return best, timings
