"""GPU parity of the nodal DC path against the FP64 CPU oracle (scipy splu).

Tolerances: the operator application in fp64 matches the assembled CSR operator to 1e-13
(relative l2); potentials match the direct solve to 1e-7, data / Jvec / Jtvec / getJ to 1e-6
(relative l2), the bar of the nodal DC target.
"""

import numpy as np
import pytest
import torch

from tests._helpers import oracle_for, rel, tiny_problem

pytestmark = pytest.mark.gpu


def _sim(mesh, survey, bc="Robin", precision="mixed", **kw):
    from simpeg_amd import maps
    from simpeg_amd.resistivity import Simulation3DNodal

    return Simulation3DNodal(mesh, survey=survey, bc_type=bc, sigmaMap=maps.ExpMap(mesh),
                             precision=precision, **kw)


@pytest.fixture(scope="module")
def tiny():
    return tiny_problem()


@pytest.mark.parametrize("bc", ["Robin", "Neumann"])
def test_apply_matches_assembled_operator(gpu, tiny, bc):
    mesh, survey, m = tiny
    sim = _sim(mesh, survey, bc)
    orc = oracle_for(mesh, survey, bc)
    orc.set_model(m)
    A = orc.getA()
    sim.model = m
    sim._ensure_sigma()
    eng = sim.engine
    rng = np.random.default_rng(1)
    x = rng.standard_normal((mesh.n_nodes, 3))
    y = eng.from_block(eng.apply(eng.to_block(x)))
    assert rel(y, A @ x) < 1e-13
    y32 = eng.from_block(eng.apply(eng.to_block(x, torch.float32)))
    assert rel(y32, A @ x) < 2e-6


@pytest.mark.parametrize("interp", [0, 1])
@pytest.mark.parametrize("wlevel,wsmooth,nuc", [(0, 0, 0), (1, 0, 0), (1, 1, 0), (1, 1, 1)])
def test_vcycle_is_symmetric(gpu, tiny, wlevel, wsmooth, nuc, interp):
    """The fp64 V-cycle (W-cycle on level 1, with two or one smoothing steps between its coarse
    corrections, and one step on the coarse levels) is a symmetric positive definite operator,
    with geometric and with operator-dependent interpolation (restriction = P^T)."""
    mesh, survey, m = tiny
    sim = _sim(mesh, survey)
    sim.model = m
    sim._ensure_sigma()
    eng = sim.engine
    assert eng.levels >= 3
    eng.set_params(wcycle_level=wlevel, interp=interp, wcycle_smooth=wsmooth, nu_coarse=nuc)
    rng = np.random.default_rng(2)
    a = rng.standard_normal((mesh.n_nodes, 2))
    b = rng.standard_normal((mesh.n_nodes, 2))
    Ma = eng.from_block(eng.vcycle(eng.to_block(a)))
    Mb = eng.from_block(eng.vcycle(eng.to_block(b)))
    lhs = np.sum(Ma * b, axis=0)
    rhs = np.sum(a * Mb, axis=0)
    assert np.allclose(lhs, rhs, rtol=1e-10, atol=0)
    # positive definite on these vectors
    assert np.all(np.sum(Ma * a, axis=0) > 0)


@pytest.mark.parametrize("bc", ["Robin", "Neumann"])
def test_fp32_vcycle_bf16_storage(gpu, tiny, bc):
    """The fp32 V-cycle with bfloat16 vector storage stays within bf16 rounding of the fp32-storage
    V-cycle and of the fp64 one, and is symmetric to that accuracy."""
    mesh, survey, m = tiny
    sim = _sim(mesh, survey, bc)
    sim.model = m
    sim._ensure_sigma()
    eng = sim.engine
    rng = np.random.default_rng(3)
    a = rng.standard_normal((mesh.n_nodes, 3))
    b = rng.standard_normal((mesh.n_nodes, 3))
    ref = eng.from_block(eng.vcycle(eng.to_block(a)))
    out = {}
    for flag in (0, 1):
        eng.set_params(vcycle_bf16=flag)
        out[flag] = eng.from_block(eng.vcycle(eng.to_block(a).float()).double())
    mb = eng.from_block(eng.vcycle(eng.to_block(b).float()).double())
    eng.set_params(vcycle_bf16=2)
    assert rel(out[0], ref) < 1e-5
    assert 1e-6 < rel(out[1], ref) < 3e-2  # bf16 storage is in effect, and close
    lhs = np.sum(out[1] * b, axis=0)
    rhs = np.sum(a * mb, axis=0)
    assert np.allclose(lhs, rhs, rtol=3e-2)
    assert np.all(np.sum(out[1] * a, axis=0) > 0)


@pytest.mark.parametrize("storage", [0, 1, 2, 3, 4])
def test_mixed_solve_with_each_vcycle_storage(gpu, tiny, storage):
    """Mixed-precision solves with fp32 multigrid vectors (0), bf16 vectors and a bf16
    preconditioned residual (1), bf16 vectors with an fp32 one (2), bf16 vectors fed from a
    bf16 copy of the residual (3), and both (4) all reach the fp64 tolerance and the direct
    solver's potentials."""
    mesh, survey, m = tiny
    sim = _sim(mesh, survey, gpu_params={"vcycle_bf16": storage})
    orc = oracle_for(mesh, survey)
    u_ref = orc.fields(m)
    f = sim.fields(m)
    assert sim.last_solve.converged and sim.last_solve.max_rel_res <= 1e-10
    # an ordinary model never needs the fp64 continuation
    assert sim.last_solve.fp64_iters == 0, sim.last_solve
    for s in range(survey.nSrc):
        assert rel(f[:, "phiSolution"][:, s], u_ref[:, s]) < 1e-7


@pytest.mark.parametrize("interp", [0, 1])
@pytest.mark.parametrize("bc", ["Robin", "Neumann"])
def test_fused_coarse_tail_matches_multi_launch_cycle(gpu, tiny, bc, interp):
    """The coarse levels run as one LDS-resident launch give the V-cycle of the separate
    launches (same smoothing, transfers and coarse inverse; fp32 in LDS instead of bf16 vectors),
    and the fused cycle stays symmetric to that accuracy."""
    mesh, survey, m = tiny
    sim = _sim(mesh, survey, bc)
    sim.model = m
    sim._ensure_sigma()
    eng = sim.engine
    assert eng.levels >= 3
    rng = np.random.default_rng(5)
    a = rng.standard_normal((mesh.n_nodes, 4))
    b = rng.standard_normal((mesh.n_nodes, 4))
    out = {}
    for flag in (0, 1):
        eng.set_params(fused_tail=flag, wcycle_level=0, interp=interp)
        out[flag] = eng.from_block(eng.vcycle(eng.to_block(a, torch.float32)).double())
    mb = eng.from_block(eng.vcycle(eng.to_block(b, torch.float32)).double())
    eng.set_params(fused_tail=1, wcycle_level=-1, interp=1)
    assert 1e-7 < rel(out[1], out[0]) < 3e-2  # the fused path ran, and agrees
    lhs = np.sum(out[1] * b, axis=0)
    rhs = np.sum(a * mb, axis=0)
    assert np.allclose(lhs, rhs, rtol=3e-2)
    assert np.all(np.sum(out[1] * a, axis=0) > 0)


@pytest.mark.parametrize("bc", ["Robin", "Neumann"])
@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_fields_dpred_match_direct_solver(gpu, tiny, bc, precision):
    mesh, survey, m = tiny
    sim = _sim(mesh, survey, bc, precision)
    orc = oracle_for(mesh, survey, bc)
    u_ref = orc.fields(m)
    f = sim.fields(m)
    u = f[:, "phiSolution"]
    for s in range(survey.nSrc):
        assert rel(u[:, s], u_ref[:, s]) < 1e-7, s
    assert sim.last_solve.converged
    d = sim.dpred(m, f=f)
    assert rel(d, orc.dpred(m, u=u_ref)) < 1e-6


@pytest.mark.parametrize("bc", ["Robin", "Neumann"])
def test_jvec_jtvec_match_oracle(gpu, tiny, bc):
    mesh, survey, m = tiny
    sim = _sim(mesh, survey, bc)
    orc = oracle_for(mesh, survey, bc)
    u_ref = orc.fields(m)
    f = sim.fields(m)
    rng = np.random.default_rng(3)
    v = rng.standard_normal(mesh.n_cells)
    w = rng.standard_normal(survey.nD)
    assert rel(sim.Jvec(m, v, f=f), orc.Jvec(m, v, u=u_ref)) < 1e-6
    assert rel(sim.Jtvec(m, w, f=f), orc.Jtvec(m, w, u=u_ref)) < 1e-6


def test_getJ_and_JtJdiag_match_oracle(gpu, tiny):
    mesh, survey, m = tiny
    sim = _sim(mesh, survey)
    orc = oracle_for(mesh, survey)
    u_ref = orc.fields(m)
    J_ref = orc.getJ(m, u=u_ref)
    J = sim.getJ(m)
    assert J.shape == J_ref.shape
    assert rel(J, J_ref) < 1e-6
    import scipy.sparse as sp

    W = sp.diags(np.linspace(0.5, 2.0, survey.nD))
    sim._gtgdiag = None
    ref = orc.getJtJdiag(m, W=W, u=u_ref)
    assert rel(sim.getJtJdiag(m, W=W), ref) < 1e-6
    # without a stored J the diagonal is accumulated on the GPU
    sim2 = _sim(mesh, survey)
    assert rel(sim2.getJtJdiag(m, W=W), ref) < 1e-6
    assert sim2._Jmatrix is None


def test_small_block_model_against_oracle(gpu):
    from simpeg_amd import workloads as W

    wl = W.small_block_model(0)
    sim = _sim(wl.mesh, wl.survey)
    orc = oracle_for(wl.mesh, wl.survey)
    u_ref = orc.fields(wl.model)
    f = sim.fields(wl.model)
    u = f[:, "phiSolution"]
    errs = [rel(u[:, s], u_ref[:, s]) for s in range(wl.survey.nSrc)]
    print("max potential rel err", max(errs), "solve", sim.last_solve)
    assert max(errs) < 1e-7
    assert rel(sim.dpred(wl.model, f=f), orc.dpred(wl.model, u=u_ref)) < 1e-6
    rng = np.random.default_rng(4)
    v = rng.standard_normal(wl.mesh.n_cells)
    w = rng.standard_normal(wl.survey.nD)
    assert rel(sim.Jvec(wl.model, v, f=f), orc.Jvec(wl.model, v, u=u_ref)) < 1e-6
    assert rel(sim.Jtvec(wl.model, w, f=f), orc.Jtvec(wl.model, w, u=u_ref)) < 1e-6


def test_small_block_model_against_golden_fixture(gpu):
    import os

    from simpeg_amd import workloads as W

    gold = np.load(os.path.join(os.path.dirname(__file__), "golden", "small_block_model.npz"))
    wl = W.small_block_model(0)
    sim = _sim(wl.mesh, wl.survey)
    f = sim.fields(wl.model)
    u = f[:, "phiSolution"]
    for k, s in enumerate(gold["phi_sources"]):
        assert rel(u[:, s], gold["phi"][:, k]) < 1e-7
    assert rel(sim.dpred(wl.model, f=f), gold["dpred"]) < 1e-6
    v = np.random.default_rng(11).standard_normal(wl.mesh.n_cells)
    w = np.random.default_rng(12).standard_normal(wl.survey.nD)
    assert rel(sim.Jvec(wl.model, v, f=f), gold["jvec"]) < 1e-6
    assert rel(sim.Jtvec(wl.model, w, f=f), gold["jtvec"]) < 1e-6


@pytest.mark.parametrize("precision", ["mixed", "mixed_ir", "fp64"])
def test_precision_modes_agree(gpu, tiny, precision):
    mesh, survey, m = tiny
    sim = _sim(mesh, survey, precision=precision)
    orc = oracle_for(mesh, survey)
    u_ref = orc.fields(m)
    f = sim.fields(m)
    assert sim.last_solve.converged
    assert sim.last_solve.max_rel_res <= (1.5e-10 if precision == "fp64" else 1e-10)
    for s in range(survey.nSrc):
        assert rel(f[:, "phiSolution"][:, s], u_ref[:, s]) < 1e-7


def test_mode_switching_on_one_handle(gpu, tiny):
    """Solves in different precision modes on one handle must not leak state (convergence
    flags, reduction scalars) into each other."""
    mesh, survey, m = tiny
    sim = _sim(mesh, survey)
    orc = oracle_for(mesh, survey)
    u_ref = orc.fields(m)
    for mode in ("mixed_ir", "mixed", "fp64", "mixed", "mixed_ir", "mixed"):
        sim.precision = mode
        f = sim.fields(m)
        assert sim.last_solve.converged, mode
        for s in range(survey.nSrc):
            assert rel(f[:, "phiSolution"][:, s], u_ref[:, s]) < 1e-7, mode


def test_miniaturize_equivalence(gpu, tiny):
    """tests/em/static/test_DC_miniaturize.py:285-312: miniaturize=True gives the same dpred,
    Jvec, Jtvec and J as the plain survey."""
    from tests._helpers import mixed_survey

    mesh, _, m = tiny
    survey = mixed_survey()
    s1 = _sim(mesh, survey)
    s2 = _sim(mesh, survey, miniaturize=True)
    assert s2._sim_survey.nSrc != survey.nSrc
    f1, f2 = s1.fields(m), s2.fields(m)
    assert np.allclose(s1.dpred(m, f=f1), s2.dpred(m, f=f2), rtol=1e-7, atol=1e-12)
    rng = np.random.default_rng(42)
    u = rng.uniform(size=m.shape)
    v = rng.uniform(size=survey.nD)
    assert rel(s2.Jvec(m, u, f=f2), s1.Jvec(m, u, f=f1)) < 1e-6
    assert rel(s2.Jtvec(m, v, f=f2), s1.Jtvec(m, v, f=f1)) < 1e-6
    assert rel(s2.getJ(m, f=f2), s1.getJ(m, f=f1)) < 1e-6
    orc = oracle_for(mesh, survey)
    assert rel(s2.dpred(m, f=f2), orc.dpred(m)) < 1e-6


def test_ip_nodal_is_scaled_dc_sensitivity(gpu, tiny):
    """IP data d = s * J_dc(sigma) eta with sigmaDeriv = -diag(sigma) etaDeriv
    (induced_polarization/simulation.py:37-39, 117-125); checked against the oracle's DC Jvec."""
    from simpeg_amd import maps
    from simpeg_amd.induced_polarization import Simulation3DNodal as IPNodal

    mesh, survey, m = tiny
    sigma = np.exp(m)
    rng = np.random.default_rng(8)
    eta = 0.05 + 0.02 * rng.random(mesh.n_cells)
    ip = IPNodal(mesh, survey=survey, sigma=sigma, etaMap=maps.IdentityMap(mesh))
    d = ip.dpred(eta)
    orc = oracle_for(mesh, survey, mapping="identity")
    assert rel(d, orc.Jvec(sigma, -sigma * eta)) < 1e-6
    v = rng.standard_normal(mesh.n_cells)
    w = rng.standard_normal(survey.nD)
    lhs = w @ ip.Jvec(eta, v)
    rhs = v @ ip.Jtvec(eta, w)
    assert abs(lhs - rhs) < 1e-8 * abs(lhs)
    assert rel(ip.Jtvec(eta, w), orc.Jtvec(sigma, w) * -sigma) < 1e-6


def test_autotuner_picks_converging_parameters(gpu, tiny, tmp_path):
    from simpeg_amd.autotune import tune

    mesh, survey, m = tiny
    sim = _sim(mesh, survey)
    sim.model = m
    sim._ensure_sigma()
    q = sim._source_block()
    cache = str(tmp_path / "tune.json")
    best, timings = tune(sim.engine, q, survey.nSrc, cache_path=cache)
    assert len(timings) >= 3 and all(t["seconds"] > 0 for t in timings)
    best2, _ = tune(sim.engine, q, survey.nSrc, cache_path=cache)
    assert best2 == best
    # a cached choice overrides only the tuned fields
    sim.engine.set_params(graph=1 - best["graph"])
    best3, _ = tune(sim.engine, q, survey.nSrc, cache_path=cache)
    assert best3["graph"] == 1 - best["graph"] and sim.engine.params.graph == best3["graph"]
    assert {k: v for k, v in best3.items() if k != "graph"} == \
        {k: v for k, v in best.items() if k != "graph"}
    orc = oracle_for(mesh, survey)
    u_ref = orc.fields(m)
    f = sim.fields(m)
    assert rel(f[:, "phiSolution"][:, 0], u_ref[:, 0]) < 1e-7


@pytest.mark.parametrize("bc", ["Robin", "Neumann"])
def test_solver_wrapper_matches_direct_factorisation(gpu, tiny, bc):
    """Ainv = NodalGPUSolver(A, mesh=mesh); Ainv * b against splu of the same assembled A, for
    the simulation's source term (several columns) and for one random vector (which, for the
    Neumann operator, has a non-zero entry in the identity row 0)."""
    from simpeg_amd.solvers import NodalGPUSolver

    mesh, survey, m = tiny
    orc = oracle_for(mesh, survey, bc)
    orc.set_model(m)
    A = orc.getA()
    Ainv = NodalGPUSolver(A, mesh=mesh)
    try:
        q = orc.source_term()
        u = Ainv * q
        u_ref = orc.solve(q)
        for s in range(q.shape[1]):
            assert rel(u[:, s], u_ref[:, s]) < 1e-7, s
        assert Ainv.last_solve.converged
        b = np.random.default_rng(4).standard_normal(mesh.n_nodes)
        x = Ainv * b
        assert x.shape == b.shape
        assert rel(x, orc.solve(b)) < 1e-7
        assert rel(A @ x, b) < 1e-9
    finally:
        Ainv.clean()


def test_graph_replay_matches_launched_cycle(gpu, tiny):
    """With gpu_params graph=1 the finest-level V-cycle is captured once per workspace and
    column count and replayed every PCG iteration, with the solver's own launches and host
    polls between replays.  Every result of a sequence of solves -- fields, Jtvec, fields for a
    new model (same graph, new coefficients), a 2-column solve (a second graph), the full batch
    again -- must equal the launched cycle's bit for bit."""
    mesh, survey, m = tiny
    rng = np.random.default_rng(3)
    v = rng.standard_normal(survey.nD)
    m2 = m + 0.2 * rng.standard_normal(m.size)
    res = []
    for g in (0, 1):
        sim = _sim(mesh, survey, gpu_params={"graph": g})
        f = sim.fields(m)
        out = [f[:, "phiSolution"].copy(), sim.Jtvec(m, v, f=f)]
        f2 = sim.fields(m2)
        u2 = f2[:, "phiSolution"].copy()
        out.append(u2)
        eng = sim.engine
        q = eng.to_block(u2[:, :2])
        u, _ = eng.solve(q, 2, mode="mixed", tol=1e-8)
        out.append(eng.from_block(u, 2))
        out.append(sim.fields(m)[:, "phiSolution"].copy())
        res.append(out)
    for a, b in zip(*res):
        assert np.array_equal(a, b)
    assert rel(res[1][0], oracle_for(mesh, survey).fields(m)) < 1e-7


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_ragged_batches_and_zero_columns(gpu, tiny, precision):
    """Column counts that are not a multiple of the columns per work-item (1, 3, 13, 33), a
    zero right-hand side among them (stays exactly zero, counts as converged), an empty block,
    and two blocks of different widths on one handle, against splu of the same operator."""
    from simpeg_amd.solvers import NodalGPUSolver

    mesh, survey, m = tiny
    orc = oracle_for(mesh, survey)
    orc.set_model(m)
    A = orc.getA()
    Ainv = NodalGPUSolver(A, mesh=mesh, precision=precision)
    rng = np.random.default_rng(11)
    try:
        assert Ainv.solve(np.zeros((mesh.n_nodes, 0))).shape == (mesh.n_nodes, 0)
        for k in (1, 3, 13, 33, 3):
            b = rng.standard_normal((mesh.n_nodes, k))
            if k > 1:
                b[:, 1] = 0.0
            x = Ainv * b
            assert Ainv.last_solve.converged
            ref = orc.solve(b)
            for c in range(k):
                if k > 1 and c == 1:
                    assert not x[:, c].any()
                else:
                    assert rel(x[:, c], ref[:, c]) < 1e-7, (k, c)
        with pytest.raises(ValueError):
            Ainv * np.ones(mesh.n_nodes + 1)
    finally:
        Ainv.clean()


def test_operator_dependent_interpolation_is_geometric_in_a_homogeneous_medium(gpu, tiny):
    """With constant sigma the two edges of a node along a grid line have conductances in the
    ratio of the inverse cell widths on every level, so the operator-dependent weights equal the
    geometric ones (the padded mesh has unequal widths, which pins the along-line indexing): the
    V-cycles agree to rounding, fp64 multi-launch and fp32 fused tail alike."""
    mesh, survey, _ = tiny
    sim = _sim(mesh, survey)
    sim.model = np.full(mesh.n_cells, np.log(0.01))
    sim._ensure_sigma()
    eng = sim.engine
    a = np.random.default_rng(8).standard_normal((mesh.n_nodes, 3))
    out64, out32 = {}, {}
    for interp in (0, 1):
        eng.set_params(interp=interp)
        out64[interp] = eng.from_block(eng.vcycle(eng.to_block(a)))
        out32[interp] = eng.from_block(eng.vcycle(eng.to_block(a, torch.float32)).double())
    assert rel(out64[1], out64[0]) < 1e-12
    assert rel(out32[1], out32[0]) < 1e-5


def test_operator_dependent_interpolation_on_a_high_contrast_model(gpu):
    """Layers of 0.005 / 0.02 / 0.002 S/m with conductive ellipsoids (0.2-0.5 S/m): the
    operator-dependent interpolation needs no more PCG iterations than the geometric one, and
    both reach the direct solver's potentials."""
    from simpeg_amd import workloads as W

    wl = W.layered_ellipsoid_model(0, core=(24, 24, 16), npad=4, nlines=1, nsrc_per_line=2,
                                   nrx=4)
    orc = oracle_for(wl.mesh, wl.survey)
    u_ref = orc.fields(wl.model)
    its = {}
    for interp in (0, 1):
        sim = _sim(wl.mesh, wl.survey, rtol=1e-8, gpu_params={"interp": interp})
        u = sim.fields(wl.model)[:, "phiSolution"]
        its[interp] = sim.last_solve.iters
        for s in range(wl.survey.nSrc):
            assert rel(u[:, s], u_ref[:, s]) < 1e-7
    assert its[1] <= its[0]


@pytest.mark.parametrize("shape", [(12, 1, 8), (3, 17, 2), (21, 9, 13)])
@pytest.mark.parametrize("bc", ["Robin", "Neumann"])
def test_thin_and_odd_meshes(gpu, shape, bc):
    """A one-cell-thick mesh, an axis with 2 cells and odd cell counts (unequal coarsening
    gaps, axes that stop coarsening early): potentials against splu for random right-hand
    sides through the solver wrapper, both interpolations."""
    from simpeg_amd.mesh import TensorMesh
    from simpeg_amd.solvers import NodalGPUSolver
    from oracle import OracleDC, OracleMesh

    rng = np.random.default_rng(sum(shape))
    h = [rng.uniform(5.0, 15.0, n) for n in shape]
    mesh = TensorMesh(h)
    om = OracleMesh(mesh.h, origin=mesh.origin)
    orc = OracleDC(om, [], bc_type=bc)
    orc.set_model(np.log(0.01) + rng.standard_normal(mesh.n_cells))
    A = orc.getA()
    b = rng.standard_normal((mesh.n_nodes, 3))
    ref = orc.solve(b)
    for interp in (0, 1):
        Ainv = NodalGPUSolver(A, mesh=mesh, gpu_params={"interp": interp})
        try:
            x = Ainv * b
            assert Ainv.last_solve.converged
            for c in range(3):
                assert rel(x[:, c], ref[:, c]) < 1e-7, (interp, c)
        finally:
            Ainv.clean()


@pytest.mark.parametrize("contrast,bc", [(1e3, "Robin"), (1e4, "Robin"), (1e5, "Robin"),
                                         (1e5, "Neumann")])
def test_extreme_conductivity_contrast(gpu, tiny, contrast, bc):
    """A conductive block of contrast 1e3 .. 1e5 against the background (and a resistive one of
    the inverse contrast): the mixed solve still reaches the fp64 tolerance and the direct
    solver's potentials, with the Robin and the Neumann (pinned node 0) operator.  From 1e4 the
    fp32 recursion stops tracking the true residual; the stall guard hands the solve to the
    fp64 continuation."""
    mesh, survey, m = tiny
    nx, ny, nz = mesh.shape_cells
    s = np.full((nx, ny, nz), 1e-2)
    s[4:8, 5:9, 3:6] = 1e-2 * contrast
    s[9:12, 3:6, 4:8] = 1e-2 / contrast
    mm = np.log(s.ravel(order="F"))
    orc = oracle_for(mesh, survey, bc)
    u_ref = orc.fields(mm)
    sim = _sim(mesh, survey, bc)
    f = sim.fields(mm)
    u = f[:, "phiSolution"]
    res = sim.last_solve
    assert res.converged and res.max_rel_res <= 1e-10, res
    if contrast >= 1e4 and bc == "Robin":
        # measured: the fp32 phase stalls on these models, so the continuation is exercised
        assert res.fp64_iters > 0, res
    for c in range(survey.nSrc):
        assert rel(u[:, c], u_ref[:, c]) < 1e-7, (contrast, c, res)
    # the adjoint solve reuses the operator: after a stall it runs fp64 from the start
    w = np.random.default_rng(3).standard_normal(survey.nD)
    jt = sim.Jtvec(mm, w, f=f)
    adj = sim.last_solve
    assert adj.converged, adj
    if res.fp64_iters:
        assert adj.iters == 0 and adj.fp64_iters > 0, adj
    assert rel(jt, orc.Jtvec(mm, w, u=u_ref)) < 1e-6


@pytest.mark.parametrize("bc", ["Robin", "Neumann"])
def test_adjoint_contraction_column_groups(gpu, tiny, bc):
    """sdc_jtvec_contract and sdc_jtj_diag_accumulate against the oracle's getADeriv adjoint
    (map derivative = identity) for column counts below, at and across the kernel's 8-column
    groups (1, 8, 9, 17), each lambda column paired with a drawn u column: the sum over columns,
    one row per column, and the W^2-weighted squares added onto existing values."""
    import scipy.sparse as sp

    mesh, survey, m = tiny
    sim = _sim(mesh, survey, bc)
    orc = oracle_for(mesh, survey, bc)
    orc.set_model(m)
    orc.sigma_deriv = sp.identity(mesh.n_cells, format="csr")
    sim.model = m
    sim._ensure_sigma()
    eng = sim.engine
    rng = np.random.default_rng(7)
    nu = 5
    U = rng.standard_normal((mesh.n_nodes, nu))
    ub = eng.to_block(U)
    for ncol in (1, 8, 9, 17):
        L = rng.standard_normal((mesh.n_nodes, ncol))
        ucol = rng.integers(0, nu, ncol)
        ref = np.stack([orc.getADeriv(U[:, ucol[j]], L[:, j], adjoint=True) for j in range(ncol)])
        lb = eng.to_block(L)
        got = eng.contract(ub, lb, ncol, ucol=ucol).cpu().numpy()
        assert rel(got, ref.sum(0)) < 1e-12, ncol
        per = eng.contract(ub, lb, ncol, ucol=ucol, per_column=True).cpu().numpy()
        assert rel(per, ref) < 1e-12, ncol
        w = rng.random(ncol)
        base = rng.random(mesh.n_cells)
        out = torch.as_tensor(base).to(eng.device)
        eng.jtj_diag_accumulate(ub, lb, ncol, ucol, w, out)
        assert rel(out.cpu().numpy(), base + (w[:, None] * ref ** 2).sum(0)) < 1e-12, ncol
