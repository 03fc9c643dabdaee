"""Stencil kernels with column loops (sdc_params.col_loop 8 / 16: each work-item takes its
columns one at a time with the node's coefficients held) against the default kernels and the
oracle's direct solver, Robin and Neumann, single operator and operator sets (where one
work-item's columns cross a set boundary)."""

import numpy as np
import pytest

from tests._helpers import oracle_for, rel, tiny_problem

pytestmark = pytest.mark.gpu


def _columns(eng, q0, nsrc, ncol):
    """ncol right-hand sides: the survey's sources scaled 1, 1.25, 1.5, ..."""
    q = eng.block(ncol)
    for j in range(ncol):
        q[j] = q0[j % nsrc] * (1.0 + 0.25 * j)
    return q


@pytest.mark.parametrize("bc", ["Robin", "Neumann"])
@pytest.mark.parametrize("mode", ["mixed", "fp64"])
@pytest.mark.parametrize("col_loop", [8, 16])
def test_column_loops_match_default_kernels_and_oracle(gpu, bc, mode, col_loop):
    from simpeg_amd import maps
    from simpeg_amd.resistivity import Simulation3DNodal

    mesh, survey, m = tiny_problem(seed=4, nsrc=4)
    sim = Simulation3DNodal(mesh, survey=survey, sigmaMap=maps.ExpMap(mesh), bc_type=bc)
    eng = sim.engine
    eng.set_sigma(np.exp(m))
    ncol = 37  # ragged: two full loops of 16 and a partial one
    q = _columns(eng, sim._source_block(), survey.nSrc, ncol)
    eng.set_params(col_loop=1)
    u1, r1 = eng.solve(q, ncol, mode=mode, tol=1e-11)
    eng.set_params(col_loop=col_loop)
    u2, r2 = eng.solve(q, ncol, mode=mode, tol=1e-11)
    assert r1.converged and r2.converged
    for j in range(ncol):
        assert rel(u2[j, :eng.nN].cpu().numpy(), u1[j, :eng.nN].cpu().numpy()) < 1e-9
    ur = oracle_for(mesh, survey, bc_type=bc).fields(m)
    for j in range(survey.nSrc):
        assert rel(u2[j, :eng.nN].cpu().numpy() / (1.0 + 0.25 * j), ur[:, j]) < 1e-8


@pytest.mark.parametrize("bc", ["Robin", "Neumann"])
def test_column_loops_cross_operator_set_boundaries(gpu, bc):
    from simpeg_amd import maps
    from simpeg_amd.resistivity import Simulation3DNodal

    mesh, survey, m = tiny_problem(seed=5, nsrc=4)
    sim = Simulation3DNodal(mesh, survey=survey, sigmaMap=maps.ExpMap(mesh), bc_type=bc)
    eng = sim.engine
    rng = np.random.default_rng(9)
    models = [m, m + 0.7 * rng.standard_normal(m.size), np.full(m.size, np.log(0.03))]
    eng.set_sigma_sets(np.stack([np.exp(mk) for mk in models]))
    colset = np.repeat(np.array([0, 1, 2, 1, 0], np.int32), 8)  # 40 columns, sets per 8
    q = _columns(eng, sim._source_block(), survey.nSrc, colset.size)
    eng.set_params(col_loop=1)
    u1, _ = eng.solve_sets(q, colset, mode="mixed", tol=1e-11)
    eng.set_params(col_loop=16)
    u2, _ = eng.solve_sets(q, colset, mode="mixed", tol=1e-11)
    for j in range(colset.size):
        assert rel(u2[j, :eng.nN].cpu().numpy(), u1[j, :eng.nN].cpu().numpy()) < 1e-9
    # set 2 against the direct solver
    ur = oracle_for(mesh, survey, bc_type=bc).fields(models[2])
    for j in range(16, 20):
        k = j % survey.nSrc
        assert rel(u2[j, :eng.nN].cpu().numpy() / (1.0 + 0.25 * j), ur[:, k]) < 1e-8
