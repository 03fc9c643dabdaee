"""The mixed PCG's ring of search directions: with a reliable-update threshold of 1e-9 (a
residual drop of nine decades before the first update), the solve runs more iterations between
updates than the ring has slots (6), so the device folds the pending directions into u without
a residual replacement.  Solutions must match the
oracle's direct solve and the default schedule, with iterations recorded as graphs and launched
one by one, one and several iterations per host read."""

import numpy as np
import pytest

from tests._helpers import oracle_for, rel, tiny_problem

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("graph", [1, 0])
@pytest.mark.parametrize("check_every", [1, 3])
@pytest.mark.parametrize("bc", ["Robin", "Neumann"])
def test_full_ring_folds_without_residual_replacement(gpu, bc, check_every, graph):
    from simpeg_amd import maps
    from simpeg_amd.resistivity import Simulation3DNodal

    mesh, survey, m = tiny_problem(seed=8, nsrc=4)
    sim = Simulation3DNodal(mesh, survey=survey, sigmaMap=maps.ExpMap(mesh), bc_type=bc)
    eng = sim.engine
    eng.set_sigma(np.exp(m))
    q = sim._source_block()
    ns = survey.nSrc
    eng.set_params(check_every=check_every, graph=graph)
    u0, r0 = eng.solve(q, ns, mode="mixed", tol=1e-10)
    # the first update after a drop of nine decades: more than 6 directions pile up before it
    eng.set_params(reliable_delta=1e-9)
    u1, r1 = eng.solve(q, ns, mode="mixed", tol=1e-10)
    eng.set_params(reliable_delta=1e-4)
    assert r0.converged and r1.converged
    assert r1.iters > 7
    ur = oracle_for(mesh, survey, bc_type=bc).fields(m)
    for j in range(ns):
        a = u1[j, :eng.nN].cpu().numpy()
        assert rel(a, ur[:, j]) < 1e-8
        assert rel(a, u0[j, :eng.nN].cpu().numpy()) < 1e-8
