"""Recorded PCG iterations (sdc_params.graph with check_every iterations per replay) serve
solves with different tolerances, iteration limits and precision modes on one handle: each
solve follows its own parameters, checked against the oracle's assembled operator."""

import numpy as np
import pytest

from tests._helpers import oracle_for, tiny_problem

pytestmark = pytest.mark.gpu


def _residual(orc, q_host, u_host):
    A = orc.getA()
    r = q_host - A @ u_host
    return np.linalg.norm(r, axis=0) / np.linalg.norm(q_host, axis=0)


@pytest.mark.parametrize("mode", ["mixed", "fp64", "mixed_ir"])
@pytest.mark.parametrize("check_every", [1, 2, 3])
def test_recorded_iterations_follow_each_solves_parameters(gpu, mode, check_every):
    from simpeg_amd import maps
    from simpeg_amd.resistivity import Simulation3DNodal

    mesh, survey, m = tiny_problem(seed=2, nsrc=4)
    sim = Simulation3DNodal(mesh, survey=survey, sigmaMap=maps.ExpMap(mesh))
    eng = sim.engine
    eng.set_params(graph=1, check_every=check_every)
    eng.set_sigma(np.exp(m))
    q = sim._source_block()
    nS = survey.nSrc
    orc = oracle_for(mesh, survey)
    orc.set_model(m)
    qh = eng.from_block(q, nS)
    iters = {}
    for tol in (1e-4, 1e-10, 1e-6, 1e-11, 1e-4):
        u, r = eng.solve(q, nS, mode=mode, tol=tol)
        assert r.converged and r.max_rel_res <= tol
        res = _residual(orc, qh, eng.from_block(u, nS))
        assert res.max() <= 1.05 * tol
        iters.setdefault(tol, []).append(r.iters + r.fp64_iters)
    # a loose solve recorded after a tight one stops as early as the first loose solve
    assert iters[1e-4][0] == iters[1e-4][1]
    assert iters[1e-4][0] < iters[1e-10][0]
