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
