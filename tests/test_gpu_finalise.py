"""The mixed PCG's per-iteration finalisers with their decisions in the last work-group: the
column norms then the reliable-update decision (k_fin_decide), the true residuals then the stop
rules (k_fin_true_update), both drawing tickets from a counter in the control block, and the
control block written to pinned host memory at the stop.  Repeated solves on one workspace
(the counter must be back at zero after each launch), with another column count in between,
must give the same bits and match the oracle's direct solve; launched one by one and recorded
as graphs, one and several iterations per host read."""

import numpy as np
import pytest

from tests._helpers import oracle_for, rel, tiny_problem

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("graph", [1, 0])
@pytest.mark.parametrize("check_every", [1, 3])
@pytest.mark.parametrize("bc", ["Robin", "Neumann"])
def test_repeated_solves_with_last_group_finalisers(gpu, bc, check_every, graph):
    from simpeg_amd import maps
    from simpeg_amd.resistivity import Simulation3DNodal

    mesh, survey, m = tiny_problem(seed=12, nsrc=6)
    sim = Simulation3DNodal(mesh, survey=survey, sigmaMap=maps.ExpMap(mesh), bc_type=bc)
    eng = sim.engine
    eng.set_sigma(np.exp(m))
    q = sim._source_block()
    ns = survey.nSrc
    eng.set_params(check_every=check_every, graph=graph)
    u1, r1 = eng.solve(q, ns, mode="mixed", tol=1e-10)
    # fewer columns: fewer work-groups draw tickets
    u3, r3 = eng.solve(q[:3].clone(), 3, mode="mixed", tol=1e-10)
    u2, r2 = eng.solve(q, ns, mode="mixed", tol=1e-10)
    assert r1.converged and r2.converged and r3.converged
    assert r1.iters == r2.iters and r1.outer == r2.outer
    assert r1.max_rel_res == r2.max_rel_res <= 1e-10
    ur = oracle_for(mesh, survey, bc_type=bc).fields(m)
    for j in range(ns):
        a = u1[j, :eng.nN].cpu().numpy()
        assert rel(a, ur[:, j]) < 1e-8
        assert np.array_equal(a, u2[j, :eng.nN].cpu().numpy())
        if j < 3:
            assert rel(u3[j, :eng.nN].cpu().numpy(), ur[:, j]) < 1e-8
