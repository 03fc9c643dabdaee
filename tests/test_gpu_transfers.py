"""The separable restriction (the residual fused with its x factor in one kernel, then the y and
z factors) against the 27-point restriction, and both against the oracle's direct solve.

Two meshes: the padded tiny mesh, whose levels coarsen x (fused residual + x pass), and a mesh
with x cells twice as wide as the y and z cells, whose first coarsening leaves x alone (there
the residual and the restriction passes stay separate kernels).  Robin and Neumann, geometric
and operator-dependent interpolation, one operator and operator sets."""

import numpy as np
import pytest

from tests._helpers import oracle_for, rel, tiny_problem

pytestmark = pytest.mark.gpu


def _wide_x_problem():
    from simpeg_amd import workloads as W
    from simpeg_amd.mesh import TensorMesh
    from simpeg_amd.resistivity.survey import Survey

    hx = [(20.0, 3, -1.3), (20.0, 6), (20.0, 3, 1.3)]
    hy = [(10.0, 3, -1.3), (10.0, 12), (10.0, 3, 1.3)]
    hz = [(10.0, 3, -1.3), (10.0, 8)]
    mesh = TensorMesh([hx, hy, hz], origin="CCN")
    rng = np.random.default_rng(3)
    m = np.log(0.01) + 0.5 * rng.standard_normal(mesh.n_cells)
    src = W.line_sources("dipole-dipole", [-40.0, 40.0, 0.0, 0.0], 10.0, 3)[:4]
    return mesh, Survey(src), m


def _problem(name):
    return tiny_problem(seed=6, nsrc=4) if name == "tiny" else _wide_x_problem()


@pytest.mark.parametrize("interp", [0, 1])
@pytest.mark.parametrize("bc", ["Robin", "Neumann"])
@pytest.mark.parametrize("name", ["tiny", "wide_x"])
def test_separable_restriction_matches_27_point(gpu, name, bc, interp):
    from simpeg_amd import maps
    from simpeg_amd.resistivity import Simulation3DNodal

    mesh, survey, m = _problem(name)
    sim = Simulation3DNodal(mesh, survey=survey, sigmaMap=maps.ExpMap(mesh), bc_type=bc)
    eng = sim.engine
    eng.set_params(interp=interp)
    eng.set_sigma(np.exp(m))
    q = sim._source_block()
    ns = survey.nSrc
    eng.set_params(xfer_separable=1)
    u1, r1 = eng.solve(q, ns, mode="mixed", tol=1e-11)
    eng.set_params(xfer_separable=0)
    u0, r0 = eng.solve(q, ns, mode="mixed", tol=1e-11)
    assert r1.converged and r0.converged
    ur = oracle_for(mesh, survey, bc_type=bc).fields(m)
    for j in range(ns):
        a = u1[j, :eng.nN].cpu().numpy()
        assert rel(a, u0[j, :eng.nN].cpu().numpy()) < 1e-9
        assert rel(a, ur[:, j]) < 1e-8


@pytest.mark.parametrize("name", ["tiny", "wide_x"])
def test_separable_restriction_with_operator_sets(gpu, name):
    from simpeg_amd import maps
    from simpeg_amd.resistivity import Simulation3DNodal

    mesh, survey, m = _problem(name)
    sim = Simulation3DNodal(mesh, survey=survey, sigmaMap=maps.ExpMap(mesh))
    eng = sim.engine
    rng = np.random.default_rng(11)
    models = [m, m + 0.6 * rng.standard_normal(m.size)]
    eng.set_sigma_sets(np.stack([np.exp(mk) for mk in models]))
    ns = survey.nSrc
    colset = np.repeat(np.array([1, 0], np.int32), 8)  # 16 columns, 8 per set
    q = eng.block(colset.size)
    q0 = sim._source_block()
    for j in range(colset.size):
        q[j] = q0[j % ns]
    eng.set_params(xfer_separable=1)
    u1, r1 = eng.solve_sets(q, colset, mode="mixed", tol=1e-11)
    eng.set_params(xfer_separable=0)
    u0, r0 = eng.solve_sets(q, colset, mode="mixed", tol=1e-11)
    assert r1.converged and r0.converged
    for j in range(colset.size):
        assert rel(u1[j, :eng.nN].cpu().numpy(), u0[j, :eng.nN].cpu().numpy()) < 1e-9
    for k, mk in enumerate(models):
        ur = oracle_for(mesh, survey).fields(mk)
        j0 = 8 * (1 - k)  # set k holds columns [8(1-k), 8(1-k)+8)
        for j in range(j0, j0 + ns):
            assert rel(u1[j, :eng.nN].cpu().numpy(), ur[:, (j - j0) % ns]) < 1e-8
