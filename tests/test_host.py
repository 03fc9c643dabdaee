"""CPU tests of the product's host-side logic against the oracle: mesh widths and ordering,
Robin boundary weights, source term, receiver projection, survey ordering, host operator."""

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import OracleMesh
from oracle.dc_nodal import robin_boundary_matrix
from simpeg_amd import maps, workloads
from simpeg_amd.mesh import (TensorMesh, edge_mass, node_interpolation, robin_diagonal,
                             robin_face_weights, unpack_widths)
from simpeg_amd.resistivity import receivers as Rx
from simpeg_amd.resistivity import sources as Src
from simpeg_amd.resistivity.survey import Survey
from tests._helpers import rel, tiny_problem


def test_unpack_widths_and_origin():
    h = unpack_widths([(10.0, 3, -1.5), (10.0, 2), 5.0, (10.0, 2, 2.0)])
    assert np.allclose(h, [33.75, 22.5, 15.0, 10, 10, 5, 20, 40])
    m = TensorMesh([[(1.0, 4)], [(2.0, 2)], [(1.0, 3)]], origin="CCN")
    assert np.allclose(m.origin, [-2.0, -2.0, -3.0])
    assert m.shape_cells == (4, 2, 3) and m.n_nodes == 5 * 3 * 4
    om = OracleMesh(m.h, origin=m.origin)
    assert np.allclose(m.nodes, om.nodes) and np.allclose(m.cell_centers, om.cell_centers)
    assert np.allclose(m.cell_volumes, om.cell_volumes)
    assert abs(m.cell_volumes.sum() - 4 * 4 * 3) < 1e-12


def test_nodal_gradient_matches_oracle():
    mesh, _, _ = tiny_problem()
    om = OracleMesh(mesh.h, origin=mesh.origin)
    assert abs(mesh.nodal_gradient - om.nodal_gradient).max() < 1e-15


@pytest.mark.parametrize("surface", [False, True])
def test_robin_weights_match_oracle_avgbc(surface):
    mesh, _, m = tiny_problem()
    om = OracleMesh(mesh.h, origin=mesh.origin)
    sf = None
    if surface:
        rng = np.random.default_rng(0)
        n_bf = om.boundary_faces.shape[0]
        sf = rng.random(n_bf) < 0.3
    AvgBC, alpha = robin_boundary_matrix(om, sf)
    beta = robin_face_weights(mesh.h, mesh.origin, sf)
    assert np.allclose(np.nan_to_num(beta["alpha"]), np.nan_to_num(alpha), rtol=1e-13, atol=0)
    sigma = np.exp(m)
    b_ref = AvgBC @ sigma
    b = robin_diagonal(mesh.h, beta, sigma)
    assert rel(b, b_ref) < 1e-13


def test_edge_mass_matches_oracle():
    mesh, _, m = tiny_problem()
    om = OracleMesh(mesh.h, origin=mesh.origin)
    s = np.exp(m)
    assert rel(edge_mass(mesh.h, s), om.get_edge_inner_product(s).diagonal()) < 1e-14
    # e^T M_e(1) e / 3 = total volume (tests/base/test_base_pde_sim.py:126-133)
    assert abs(edge_mass(mesh.h).sum() / 3 - mesh.cell_volumes.sum()) < 1e-9 * mesh.cell_volumes.sum()


def test_interpolation_matches_oracle():
    mesh, _, _ = tiny_problem()
    om = OracleMesh(mesh.h, origin=mesh.origin)
    rng = np.random.default_rng(1)
    lo = np.array([mesh.nodes_x[0], mesh.nodes_y[0], mesh.nodes_z[0]])
    hi = np.array([mesh.nodes_x[-1], mesh.nodes_y[-1], mesh.nodes_z[-1]])
    pts = lo + rng.random((50, 3)) * (hi - lo)
    pts[:5] = mesh.nodes[[0, 7, 100, 500, mesh.n_nodes - 1]]
    idx, w = node_interpolation(mesh.h, mesh.origin, pts)
    P = sp.csr_matrix((w.ravel(), (np.repeat(np.arange(len(pts)), 8), idx.ravel())),
                      shape=(len(pts), mesh.n_nodes))
    assert abs(P - om.get_interpolation_matrix(pts)).max() < 1e-14
    # a node is reproduced exactly
    assert np.allclose(P[:5].toarray().max(axis=1), 1.0)
    with pytest.raises(ValueError):
        node_interpolation(mesh.h, mesh.origin, [[1e6, 0, 0]])


def test_host_getA_matches_oracle():
    from simpeg_amd.resistivity import Simulation3DNodal
    from tests._helpers import oracle_for

    mesh, survey, m = tiny_problem()
    for bc in ("Robin", "Neumann"):
        sim = Simulation3DNodal(mesh, survey=survey, bc_type=bc, sigmaMap=maps.ExpMap(mesh))
        sim.model = m
        orc = oracle_for(mesh, survey, bc)
        orc.set_model(m)
        A = sim.getA()
        assert abs(A - orc.getA()).max() / abs(orc.getA()).max() < 1e-13


def test_source_term_matches_oracle():
    from simpeg_amd.resistivity import Simulation3DNodal
    from tests._helpers import oracle_for

    mesh, survey, m = tiny_problem()
    sim = Simulation3DNodal(mesh, survey=survey, sigmaMap=maps.ExpMap(mesh))
    orc = oracle_for(mesh, survey)
    assert np.allclose(sim.getSourceTerm(), orc.source_term(), atol=1e-15)


def test_survey_ordering_and_generator():
    src = workloads.line_sources("dipole-dipole", [-120.0, 120.0, 40.0, 40.0], 20.0, 6)
    # 13 stations -> 10 sources, receivers n = 1..6 then 5, 4, 3, 2, 1
    assert len(src) == 10
    assert [s.nD for s in src] == [6, 6, 6, 6, 6, 5, 4, 3, 2, 1]
    sv = Survey(src)
    assert sv.nD == 45 and sv.nSrc == 10
    sl = sv.get_all_slices()
    assert sl[(src[1], src[1].receiver_list[0])] == slice(6, 12)
    assert sv.get_slice(src[9], src[9].receiver_list[0]) == slice(44, 45)
    # pole-dipole and pole-pole layouts
    pd = workloads.line_sources("pole-dipole", [0.0, 100.0, 0.0, 0.0], 10.0, 3)
    assert isinstance(pd[0], Src.Pole) and isinstance(pd[0].receiver_list[0], Rx.Dipole)
    pp = workloads.line_sources("pole-pole", [0.0, 100.0, 0.0, 0.0], 10.0, 3)
    assert isinstance(pp[0].receiver_list[0], Rx.Pole)


def test_geometric_factor_closed_form():
    a, b = np.r_[0.0, 0, 0], np.r_[10.0, 0, 0]
    M, N = np.array([[20.0, 0, 0]]), np.array([[30.0, 0, 0]])
    rx = Rx.Dipole(M, N, data_type="apparent_resistivity")
    src = Src.Dipole([rx], a, b)
    sv = Survey([src])
    G = sv.set_geometric_factor()
    ref = (1 / 20 - 1 / 10 - 1 / 30 + 1 / 20) / (2 * np.pi)
    assert np.allclose(G, ref)
    assert np.allclose(rx.geometric_factor[src], ref)


def test_api_errors():
    mesh, survey, _ = tiny_problem()
    from simpeg_amd.resistivity import Simulation3DNodal

    with pytest.raises(ValueError):
        Simulation3DNodal(mesh, survey=survey, bc_type="Dirichlet")
    with pytest.raises(TypeError):
        Src.Dipole([], location_a=[0, 0, 0])
    with pytest.raises(ValueError):
        Rx.Dipole(locations_m=np.zeros((2, 3)))
    with pytest.raises(ValueError):
        Src.Pole([], np.zeros((2, 3)))
    sim = Simulation3DNodal(mesh, survey=survey, sigmaMap=maps.ExpMap(mesh))
    with pytest.raises(ValueError):
        sim.model = np.zeros(mesh.n_cells + 1)
    with pytest.raises(AttributeError):
        Simulation3DNodal(mesh).survey


def test_grf_statistics():
    mesh = workloads.padded_mesh((24, 24, 12), npad=5)
    f = workloads.gaussian_random_field(mesh, std=1.0, corr=60.0, seed=3)
    assert f.shape == (mesh.n_cells,)
    assert 0.5 < f.std() < 1.5
    f2 = workloads.gaussian_random_field(mesh, std=1.0, corr=60.0, seed=3)
    assert np.array_equal(f, f2)


def test_miniaturize_recombination_with_oracle():
    """d = AM - AN - BM + BN from the unique pole-pole pairs reproduces the full survey."""
    from simpeg_amd.resistivity.utils import _mini_pole_pole, mini_survey_data, mini_survey_dataT
    from tests._helpers import mixed_survey, oracle_for

    mesh, _, m = tiny_problem()
    survey = mixed_survey()
    T, mini = _mini_pole_pole(survey)
    assert mini.nSrc < survey.nSrc + 10 and all(isinstance(s, Src.Pole) for s in mini.source_list)
    full = oracle_for(mesh, survey)
    small = oracle_for(mesh, mini)
    d_full = full.dpred(m)
    d_mini = small.dpred(m)
    assert np.allclose(mini_survey_data(T, d_mini), d_full, rtol=1e-10, atol=1e-14)
    # adjoint of the recombination
    rng = np.random.default_rng(0)
    a = rng.standard_normal(mini.nD)
    b = rng.standard_normal(survey.nD)
    lhs = b @ mini_survey_data(T, a)
    rhs = a @ mini_survey_dataT(T, b)
    assert np.isclose(lhs, rhs, rtol=1e-12)


@pytest.mark.parametrize("bc", ["Robin", "Neumann"])
def test_solver_reads_edge_conductances_from_assembled_operator(bc):
    """NodalGPUSolver's structure extraction: the +x/+y/+z conductances read from getA's
    sub-diagonals are the oracle's edge inner product over h^2 (G = diag(1/h) D), and the rest
    of the diagonal is AvgBC @ sigma (Robin) or zero (Neumann)."""
    from simpeg_amd.solvers import nodal_coefficients
    from tests._helpers import oracle_for

    mesh, survey, m = tiny_problem()
    orc = oracle_for(mesh, survey, bc)
    orc.set_model(m)
    A = orc.getA()
    cx, cy, cz, bd, neumann = nodal_coefficients(A, mesh)
    assert neumann == (bc == "Neumann")
    hx, hy, hz = (np.asarray(h) for h in mesh.h)
    nx, ny, nz = mesh.shape_cells
    n1x, n1y, n1z = nx + 1, ny + 1, nz + 1
    me = orc.mesh.get_edge_inner_product(orc.sigma).diagonal()
    nex, ney = nx * n1y * n1z, n1x * ny * n1z
    mex = me[:nex].reshape((nx, n1y, n1z), order="F") / hx[:, None, None] ** 2
    mey = me[nex:nex + ney].reshape((n1x, ny, n1z), order="F") / hy[None, :, None] ** 2
    mez = me[nex + ney:].reshape((n1x, n1y, nz), order="F") / hz[None, None, :] ** 2
    gx = cx.reshape((n1x, n1y, n1z), order="F")
    gy = cy.reshape((n1x, n1y, n1z), order="F")
    gz = cz.reshape((n1x, n1y, n1z), order="F")
    assert rel(gx[:nx], mex) < 1e-13 and not gx[nx].any()
    assert rel(gy[:, :ny], mey) < 1e-13 and not gy[:, ny].any()
    assert rel(gz[:, :, :nz], mez) < 1e-13 and not gz[:, :, nz].any()
    if bc == "Robin":
        assert rel(bd, orc.AvgBC @ orc.sigma) < 1e-12
    else:
        assert np.abs(bd).max() < 1e-12 * np.abs(A.diagonal()).max()


def test_solver_rejects_operators_that_are_not_the_nodal_stencil():
    from simpeg_amd.solvers import nodal_coefficients
    from tests._helpers import oracle_for

    mesh, survey, m = tiny_problem()
    orc = oracle_for(mesh, survey)
    orc.set_model(m)
    A = orc.getA().tolil()
    A[5, 200] = -1.0  # a coupling outside the 7-point pattern
    with pytest.raises(ValueError, match="7-point"):
        nodal_coefficients(A.tocsr(), mesh)
    A = orc.getA().tolil()
    A[3, 4] *= 2.0  # nonsymmetric edge
    with pytest.raises(ValueError, match="7-point"):
        nodal_coefficients(A.tocsr(), mesh)
    small = TensorMesh([[(10.0, 4)], [(10.0, 4)], [(10.0, 4)]])
    with pytest.raises(ValueError, match="nodes"):
        nodal_coefficients(orc.getA(), small)


def test_meta_adjoint_sum_matches_sparse_transpose():
    """MetaSimulation.Jtvec's accumulation equals sum_i P_i^T part_i for disjoint projections,
    a projection that selects an entry twice, and a non-projection map."""
    from simpeg_amd.meta import sum_adjoint

    rng = np.random.default_rng(0)
    n = 50
    model = rng.standard_normal(n)
    mps = [maps.Projection(n, slice(0, 20)), maps.Projection(n, slice(20, 50)),
           maps.Projection(n, np.array([3, 7, 3, 49])), maps.ExpMap(nP=n)]
    assert [mp.unique_index for mp in mps[:3]] == [True, True, False]
    parts = [rng.standard_normal(mp.shape[0]) for mp in mps]
    ref = sum(np.asarray(mp.deriv(model).T @ p).ravel() for mp, p in zip(mps, parts))
    assert np.allclose(sum_adjoint(mps, model, parts), ref, rtol=1e-14, atol=1e-14)
    # contiguous projections that tile the model (in any order) are written, not added; a gap
    # or an overlap takes the accumulating path
    for tiles in ([slice(20, 50), slice(0, 20)], [slice(0, 20), slice(21, 50)],
                  [slice(0, 25), slice(20, 50)]):
        mps = [maps.Projection(n, t) for t in tiles]
        parts = [rng.standard_normal(mp.shape[0]) for mp in mps]
        ref = sum(np.asarray(mp.deriv(model).T @ p).ravel() for mp, p in zip(mps, parts))
        assert np.array_equal(sum_adjoint(mps, model, parts), ref)


def test_edge_cell_weights_and_mesigma_deriv_match_oracle():
    """A_v equals the oracle's edge-cell averaging (so A_v sigma = edge_mass), MeSigmaDeriv
    applies diag(u) A_v sigmaDeriv forward and transposed (base/pde_simulation.py:296-321), and
    getRHS is the source term."""
    from simpeg_amd.mesh import edge_cell_weights
    from simpeg_amd.resistivity import Simulation3DNodal

    mesh, survey, m = tiny_problem()
    om = OracleMesh(mesh.h, origin=mesh.origin)
    Av = edge_cell_weights(mesh.h)
    assert abs(Av - om.edge_cell_weights).max() < 1e-12 * abs(Av).max()
    s = np.exp(m)
    assert rel(Av @ s, edge_mass(mesh.h, s)) < 1e-14
    sim = Simulation3DNodal(mesh, survey=survey, sigmaMap=maps.ExpMap(mesh))
    sim.model = m
    rng = np.random.default_rng(7)
    u = rng.standard_normal(Av.shape[0])
    v = rng.standard_normal(mesh.n_cells)
    w = rng.standard_normal(Av.shape[0])
    ref = u * (om.edge_cell_weights @ (s * v))
    assert rel(sim.MeSigmaDeriv(u, v), ref) < 1e-13
    assert rel(sim.MeSigmaDeriv(u) @ v, ref) < 1e-13
    assert np.isclose(w @ sim.MeSigmaDeriv(u, v), v @ sim.MeSigmaDeriv(u, w, adjoint=True),
                      rtol=1e-12)
    assert np.array_equal(sim.getRHS(), sim.getSourceTerm())


def test_map_derivative_transpose_paths():
    """sigmaDeriv^T g through the diagonal fast path (exp and rho maps, cached) and through the
    IP simulation's own sigmaDeriv equals the sparse transpose product."""
    from simpeg_amd.induced_polarization import Simulation3DNodal as IPNodal
    from simpeg_amd.resistivity import Simulation3DNodal

    mesh, survey, m = tiny_problem()
    g = np.random.default_rng(2).random(mesh.n_cells)
    sims = [Simulation3DNodal(mesh, survey=survey, sigmaMap=maps.ExpMap(mesh)),
            Simulation3DNodal(mesh, survey=survey, rhoMap=maps.IdentityMap(mesh)),
            IPNodal(mesh, survey=survey, sigma=np.exp(m), etaMap=maps.IdentityMap(mesh))]
    models = [m, np.exp(-m), np.full(mesh.n_cells, 0.01)]
    for sim, mm in zip(sims, models):
        sim.model = mm
        for _ in range(2):  # second call: cached derivative and diagonal flag
            assert np.allclose(sim._sigma_deriv_T(g), sim.sigmaDeriv.T @ g, rtol=1e-14)
