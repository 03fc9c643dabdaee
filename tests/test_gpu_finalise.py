"""The mixed PCG's device-side control block as the host sees it: the device writes it into
pinned host memory when the solve stops (the host reads the stop flag there after each block of
iterations, no copy per read).  Repeated solves on one workspace, with another column count in
between, must stop at the same iteration with the same statistics and bits, and match the
oracle's direct solve; launched one by one and recorded as graphs, one and several iterations
per host read."""

import numpy as np
import pytest

from tests._helpers import oracle_for, rel, tiny_problem

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("graph", [1, 0])
@pytest.mark.parametrize("check_every", [1, 3])
@pytest.mark.parametrize("bc", ["Robin", "Neumann"])
def test_repeated_solves_read_the_stop_block_from_host_memory(gpu, bc, check_every, graph):
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
    # fewer columns in between
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


@pytest.mark.parametrize("bc", ["Robin", "Neumann"])
def test_coarsest_inverse_in_global_memory(gpu, bc):
    """A coarsest level of more than 99 nodes: its Gauss-Jordan inverse runs in global memory
    instead of LDS; the solution must match the default hierarchy's and the oracle's direct
    solve.  With coarse_max 1000 the 2,550-node mesh coarsens until at most 1,000 nodes are
    left, and one coarsening step removes at most 7 of 8 nodes, so more than 125 remain."""
    from simpeg_amd import maps
    from simpeg_amd.resistivity import Simulation3DNodal

    mesh, survey, m = tiny_problem(seed=14, nsrc=4)
    sims = [Simulation3DNodal(mesh, survey=survey, sigmaMap=maps.ExpMap(mesh), bc_type=bc,
                              coarse_max=cm) for cm in (0, 1000)]
    us = []
    for sim in sims:
        eng = sim.engine
        eng.set_sigma(np.exp(m))
        u, r = eng.solve(sim._source_block(), survey.nSrc, mode="mixed", tol=1e-10)
        assert r.converged
        us.append(u)
    assert mesh.n_nodes == 2550 and 2 <= sims[1].engine.levels < sims[0].engine.levels
    ur = oracle_for(mesh, survey, bc_type=bc).fields(m)
    for j in range(survey.nSrc):
        a = us[1][j, :sims[1].engine.nN].cpu().numpy()
        assert rel(a, ur[:, j]) < 1e-8
        assert rel(a, us[0][j, :sims[0].engine.nN].cpu().numpy()) < 1e-8
