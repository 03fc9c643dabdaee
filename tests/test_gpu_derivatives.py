"""GPU parity of the derivative entry points outside Jvec/Jtvec's default path: getADeriv with
the reference's argument shapes, resistivity (rhoMap) models, whose reference derivative leaves
out the Robin boundary term, and the storeJ branches of Jvec/Jtvec."""

import numpy as np
import pytest

from tests._helpers import oracle_for, rel, tiny_problem

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tiny():
    return tiny_problem()


@pytest.mark.parametrize("bc", ["Robin", "Neumann"])
def test_getADeriv_shapes_match_oracle(gpu, tiny, bc):
    """getADeriv(u, v) forward and adjoint (resistivity/simulation.py:557-571): 1-D u with 1-D
    and 2-D v (the getJ path, _Jtvec with v=None, passes one u and a block of adjoint
    columns), and 2-D u with 2-D v (summed over the pairs), against the oracle's getADeriv."""
    from simpeg_amd import maps
    from simpeg_amd.resistivity import Simulation3DNodal

    mesh, survey, m = tiny
    sim = Simulation3DNodal(mesh, survey=survey, sigmaMap=maps.ExpMap(mesh), bc_type=bc)
    sim.model = m
    orc = oracle_for(mesh, survey, bc_type=bc)
    orc.set_model(m)
    rng = np.random.default_rng(3)
    nN, nC = mesh.n_nodes, mesh.n_cells
    u = rng.standard_normal(nN)
    v = rng.standard_normal(nC)
    assert rel(sim.getADeriv(u, v), orc.getADeriv(u, v)) < 1e-12
    lam1 = rng.standard_normal(nN)
    g1 = sim.getADeriv(u, lam1, adjoint=True)
    assert g1.shape == (nC,)
    assert rel(g1, orc.getADeriv(u, lam1, adjoint=True)) < 1e-12
    for k in (1, 3, 9, 17):
        lam = rng.standard_normal((nN, k))
        g = sim.getADeriv(u, lam, adjoint=True)
        ref = orc.getADeriv(u, lam, adjoint=True)
        if k == 1:
            assert g.shape == (nC,) and rel(g, np.ravel(ref)) < 1e-12
        else:
            assert g.shape == (nC, k)
            assert rel(g, ref) < 1e-12
    U = rng.standard_normal((nN, 5))
    L = rng.standard_normal((nN, 5))
    summed = sum(orc.getADeriv(U[:, i], L[:, i], adjoint=True) for i in range(5))
    assert rel(sim.getADeriv(U, L, adjoint=True), summed) < 1e-12
    with pytest.raises(ValueError):
        sim.getADeriv(U, L[:, :3], adjoint=True)


@pytest.mark.parametrize("bc", ["Robin", "Neumann"])
def test_rho_model_matches_reference_derivative(gpu, tiny, bc):
    """rhoMap = IdentityMap: data, Jvec and Jtvec against the oracle's resistivity model, whose
    getADeriv has no Robin term (resistivity/simulation.py:566-570); exact_bc_deriv=True
    restores it and then differs from the reference by that term."""
    from simpeg_amd import maps
    from simpeg_amd.resistivity import Simulation3DNodal

    mesh, survey, m = tiny
    rho = np.exp(-m)
    sim = Simulation3DNodal(mesh, survey=survey, rhoMap=maps.IdentityMap(mesh), bc_type=bc,
                            rtol=1e-12)
    orc = oracle_for(mesh, survey, bc_type=bc, mapping="identity", param="rho")
    u = orc.fields(rho)
    rng = np.random.default_rng(8)
    v = rho * rng.standard_normal(mesh.n_cells)
    w = rng.standard_normal(survey.nD)
    f = sim.fields(rho)
    assert rel(sim.dpred(rho, f=f), orc.dpred(rho, u=u)) < 1e-7
    jv_ref = orc.Jvec(rho, v, u=u)
    jtw_ref = orc.Jtvec(rho, w, u=u)
    assert rel(sim.Jvec(rho, v, f=f), jv_ref) < 1e-6
    assert rel(sim.Jtvec(rho, w, f=f), jtw_ref) < 1e-6
    ex = Simulation3DNodal(mesh, survey=survey, rhoMap=maps.IdentityMap(mesh), bc_type=bc,
                           rtol=1e-12, exact_bc_deriv=True)
    fe = ex.fields(rho)
    gap = rel(ex.Jtvec(rho, w, f=fe), jtw_ref)
    if bc == "Robin":
        assert gap > 1e-4
    else:
        assert gap < 1e-6


@pytest.mark.parametrize("bc", ["Robin", "Neumann"])
def test_storeJ_jvec_jtvec_match_oracle(gpu, tiny, bc):
    """storeJ=True: Jvec and Jtvec go through the stored sensitivity (resistivity/
    simulation.py:165-167, 198-200) and match the oracle and the matrix-free products."""
    from simpeg_amd import maps
    from simpeg_amd.resistivity import Simulation3DNodal

    mesh, survey, m = tiny
    sim = Simulation3DNodal(mesh, survey=survey, sigmaMap=maps.ExpMap(mesh), bc_type=bc,
                            storeJ=True, rtol=1e-12)
    free = Simulation3DNodal(mesh, survey=survey, sigmaMap=maps.ExpMap(mesh), bc_type=bc,
                             rtol=1e-12)
    orc = oracle_for(mesh, survey, bc_type=bc)
    u = orc.fields(m)
    rng = np.random.default_rng(12)
    v = rng.standard_normal(mesh.n_cells)
    w = rng.standard_normal(survey.nD)
    f = sim.fields(m)
    jv = sim.Jvec(m, v, f=f)
    assert sim._Jmatrix is not None
    assert rel(jv, orc.Jvec(m, v, u=u)) < 1e-6
    assert rel(sim.Jtvec(m, w, f=f), orc.Jtvec(m, w, u=u)) < 1e-6
    ff = free.fields(m)
    assert rel(jv, free.Jvec(m, v, f=ff)) < 1e-8
    assert rel(sim.Jtvec(m, w, f=f), free.Jtvec(m, w, f=ff)) < 1e-8
    # a new model drops the stored J
    sim.model = m + 0.1
    assert sim._Jmatrix is None
