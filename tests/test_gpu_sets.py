"""Operator sets: several conductivity models on one mesh solved as one batched solve, each
column group against its own operator (sdc_set_sigma_sets / sdc_solve_sets), checked column by
column against single-operator solves of the same right-hand sides and against the oracle's
direct solver."""

import numpy as np
import pytest
import torch

from tests._helpers import oracle_for, rel, tiny_problem

pytestmark = pytest.mark.gpu


def _padded(counts):
    """Column offsets with every problem's block padded to a multiple of 8, and the set map."""
    offs, colset = [], []
    c = 0
    for k, n in enumerate(counts):
        offs.append(c)
        w = (n + 7) // 8 * 8
        colset += [k] * w
        c += w
    return offs, np.array(colset, dtype=np.int32)


@pytest.mark.parametrize("bc", ["Robin", "Neumann"])
@pytest.mark.parametrize("mode", ["mixed", "fp64"])
def test_operator_sets_match_single_operator_solves(gpu, bc, mode):
    from simpeg_amd import maps
    from simpeg_amd.resistivity import Simulation3DNodal

    mesh, survey, m0 = tiny_problem(seed=3, nsrc=4)
    rng = np.random.default_rng(7)
    models = [m0, m0 + 0.8 * rng.standard_normal(m0.size), np.full(m0.size, np.log(0.02))]
    sims = [Simulation3DNodal(mesh, survey=survey, sigmaMap=maps.ExpMap(mesh), bc_type=bc)
            for _ in models]
    counts = [survey.nSrc, 3, survey.nSrc + 5]
    offs, colset = _padded(counts)
    eng = sims[0].engine
    ncol = colset.size
    q0 = sims[0]._source_block()
    q = eng.block(ncol)
    # right-hand sides: the survey's sources, and for the other problems shifted copies of them
    for k, n in enumerate(counts):
        for j in range(n):
            q[offs[k] + j] = q0[j % survey.nSrc] * (1.0 + 0.25 * j)
    eng.set_sigma_sets(np.stack([np.exp(mk) for mk in models]))
    u, res = eng.solve_sets(q, colset, ncol, mode=mode, tol=1e-12)
    assert res.converged and res.max_rel_res <= 1e-12
    # padding columns have a zero right-hand side and a zero solution
    for k, n in enumerate(counts):
        pad = (n + 7) // 8 * 8
        if pad > n:
            assert torch.count_nonzero(u[offs[k] + n:offs[k] + pad, :eng.nN]) == 0
    # each problem against a single-operator solve on its own engine
    for k, n in enumerate(counts):
        ek = sims[k].engine
        ek.set_sigma(np.exp(models[k]))
        qk = q[offs[k]:offs[k] + n].clone()
        uk, _ = ek.solve(qk, n, mode=mode, tol=1e-12)
        for j in range(n):
            assert rel(u[offs[k] + j, :eng.nN].cpu().numpy(), uk[j, :ek.nN].cpu().numpy()) < 1e-9
    # and one problem against the direct solver
    orc = oracle_for(mesh, survey, bc_type=bc)
    ur = orc.fields(models[2])
    for j in range(survey.nSrc):
        got = u[offs[2] + j, :eng.nN].cpu().numpy() / (1.0 + 0.25 * j)
        assert rel(got, ur[:, j]) < 1e-8


def test_operator_sets_reject_mixed_groups_and_unknown_sets(gpu):
    from simpeg_amd import maps
    from simpeg_amd.resistivity import Simulation3DNodal

    mesh, survey, m0 = tiny_problem(seed=1, nsrc=2)
    sim = Simulation3DNodal(mesh, survey=survey, sigmaMap=maps.ExpMap(mesh))
    eng = sim.engine
    eng.set_sigma_sets(np.stack([np.exp(m0), np.exp(m0 + 0.1)]))
    q = eng.block(16)
    bad = np.zeros(16, np.int32)
    bad[3] = 1  # a group of 8 columns with two sets
    with pytest.raises(ValueError, match="group of 8"):
        eng.solve_sets(q, bad, 16)
    with pytest.raises(ValueError, match="operator sets"):
        eng.solve_sets(q, np.full(16, 2, np.int32), 16)
