"""GPU parity on the benchmark workloads themselves, with checks that do not go through the GPU
stencil: the residual |q - A u| / |q| of every checked column with A assembled by the oracle
(scipy CSR, applied on the CPU) and q from the oracle's source term, agreement with the FP64
direct solver where it can factor the mesh (golden fixture of the 54,675-node mesh built like
the 932,137-node benchmark, and the smallest problems of the 48-problem batch), mixed against
fp64 GPU potentials, and the adjoint identity w^T J v = v^T J^T w.

Workloads: grf_32x32x20_32src (fixture), grf_96x96x64_128src (932,137 nodes, 128 sources),
uneven_batch_48 (48 problems of 18,513 to 356,445 nodes, 2,562 sources) and
layered_160x160x96_256src (3,289,545 nodes, 256 sources, Gauss-Newton)."""

import hashlib
import os

import numpy as np
import pytest

from tests._helpers import oracle_for, rel

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "medium_grf_model.npz")


def _oracle_operator(wl):
    """(oracle, A, q): the oracle of the workload with its model set, its assembled operator and
    its source term (nN, nSrc)."""
    orc = oracle_for(wl.mesh, wl.survey)
    orc.set_model(wl.model)
    return orc, orc.getA(), orc.source_term()


def _host_cols(sim, f, cols):
    """Potentials of the given source columns, (nN, len(cols)), copied from the device block."""
    u = f.u_block[list(cols), : sim._nN]
    return u.T.contiguous().cpu().numpy()


def _residuals(A, q, u):
    r = q - A @ u
    return np.linalg.norm(r, axis=0) / np.linalg.norm(q, axis=0)


def test_medium_grf_against_golden_fixture(gpu):
    """54,675-node Gaussian-random-field mesh (the construction of the 932k-node benchmark at the
    largest size splu factors in seconds): potentials, data, Jvec and Jtvec against the FP64
    direct solver's, stored by tests/golden/make_golden.py."""
    from simpeg_amd import maps, workloads
    from simpeg_amd.resistivity import Simulation3DNodal

    gold = np.load(GOLD)
    wl = workloads.medium_grf_model(5)
    digest = hashlib.sha256(np.ascontiguousarray(wl.model).tobytes()).digest()
    assert bytes(gold["model"].tobytes()) == digest
    sim = Simulation3DNodal(wl.mesh, survey=wl.survey, sigmaMap=maps.ExpMap(wl.mesh))
    f = sim.fields(wl.model)
    u = _host_cols(sim, f, gold["phi_sources"])
    for k in range(len(gold["phi_sources"])):
        assert rel(u[:, k], gold["phi"][:, k]) < 1e-7
    assert rel(sim.dpred(wl.model, f=f), gold["dpred"]) < 1e-6
    v = np.random.default_rng(11).standard_normal(wl.mesh.n_cells)
    w = np.random.default_rng(12).standard_normal(wl.survey.nD)
    assert rel(sim.Jvec(wl.model, v, f=f), gold["jvec"]) < 1e-6
    assert rel(sim.Jtvec(wl.model, w, f=f), gold["jtvec"]) < 1e-6


def test_grf_96x96x64_128src_residuals_fp64_and_adjoint(gpu):
    """The 932,137-node benchmark at its own tolerance (1e-8): every column's residual against
    the oracle's assembled A, the mixed potentials against fp64 GPU potentials at 1e-12, and the
    adjoint identity of Jvec / Jtvec."""
    from simpeg_amd import maps, workloads
    from simpeg_amd.resistivity import Simulation3DNodal

    wl = workloads.large_grf_model(0)
    orc, A, q = _oracle_operator(wl)
    sim = Simulation3DNodal(wl.mesh, survey=wl.survey, sigmaMap=maps.ExpMap(wl.mesh), rtol=1e-8)
    f = sim.fields(wl.model)
    print("mixed solve", sim.last_solve, flush=True)
    u = _host_cols(sim, f, range(wl.survey.nSrc))
    res = _residuals(A, q, u)
    print("max column residual", res.max(), flush=True)
    assert res.max() <= 1.05e-8
    ref = Simulation3DNodal(wl.mesh, survey=wl.survey, sigmaMap=maps.ExpMap(wl.mesh),
                            precision="fp64", rtol=1e-12)
    fr = ref.fields(wl.model)
    ur = _host_cols(ref, fr, range(wl.survey.nSrc))
    assert _residuals(A, q, ur).max() <= 1.05e-12
    errs = [rel(u[:, s], ur[:, s]) for s in range(wl.survey.nSrc)]
    print("max potential error against fp64", max(errs), flush=True)
    assert max(errs) < 1e-7
    assert rel(sim.dpred(wl.model, f=f), ref.dpred(wl.model, f=fr)) < 1e-6
    rng = np.random.default_rng(2)
    w = rng.standard_normal(wl.survey.nD)
    v = rng.standard_normal(wl.mesh.n_cells)
    assert rel(sim.Jtvec(wl.model, w, f=f), ref.Jtvec(wl.model, w, f=fr)) < 1e-6
    sim.rtol = 1e-10
    wjv = w @ sim.Jvec(wl.model, v, f=f)
    vjtw = v @ sim.Jtvec(wl.model, w, f=f)
    assert abs(wjv - vjtw) < 1e-8 * max(abs(wjv), abs(vjtw))


def test_uneven_batch_48_residuals_and_direct_solver(gpu):
    """The 48-problem batch behind MetaSimulation at the workload's settings (autotuned multigrid,
    8 streams): every column of every problem against its oracle operator; the first problems
    of sides 24 and 32 against splu potentials and data; the batch Jtvec equals the sum of the
    per-problem ones through the mappings."""
    from simpeg_amd import maps, workloads
    from simpeg_amd.meta import MetaSimulation
    from simpeg_amd.resistivity import Simulation3DNodal

    batch = workloads.uneven_batch(0)
    sizes = [wl.mesh.n_cells for wl in batch]
    offs = np.cumsum([0] + sizes)
    sims, maps_ = [], []
    for i, wl in enumerate(batch):
        sims.append(Simulation3DNodal(wl.mesh, survey=wl.survey, sigmaMap=maps.ExpMap(wl.mesh),
                                      rtol=1e-10))
        maps_.append(maps.Projection(int(offs[-1]), slice(int(offs[i]), int(offs[i + 1]))))
    meta = MetaSimulation(sims, maps_)
    m = np.concatenate([wl.model for wl in batch])
    f = meta.fields(m)
    worst = 0.0
    for i, wl in enumerate(batch):
        orc, A, q = _oracle_operator(wl)
        u = _host_cols(sims[i], f[i], range(wl.survey.nSrc))
        worst = max(worst, _residuals(A, q, u).max())
    print("worst column residual over 48 problems", worst, flush=True)
    assert worst <= 1.05e-10
    d = meta.dpred(m, f=f)
    checked = set()
    for i, wl in enumerate(batch):
        side = wl.info["side"]
        if side not in (24, 32) or side in checked:
            continue
        checked.add(side)
        orc = oracle_for(wl.mesh, wl.survey)
        ur = orc.fields(wl.model)
        u = _host_cols(sims[i], f[i], range(wl.survey.nSrc))
        assert max(rel(u[:, s], ur[:, s]) for s in range(wl.survey.nSrc)) < 1e-7
        di = d[meta._data_offsets[i]:meta._data_offsets[i + 1]]
        assert rel(di, orc.dpred(wl.model, u=ur)) < 1e-6
    assert checked == {24, 32}
    w = np.random.default_rng(4).standard_normal(meta.nD)
    g = meta.Jtvec(m, w, f=f)
    parts = [sims[i].Jtvec(batch[i].model, w[meta._data_offsets[i]:meta._data_offsets[i + 1]],
                           f=f[i]) for i in range(len(batch))]
    # the batch solves same-mesh problems together (operator sets), each problem's own Jtvec on
    # its own engine: two adjoint solves to 1e-10
    assert rel(g, np.concatenate(parts)) < 1e-8


def test_layered_160x160x96_256src_forward_adjoint_and_gauss_newton(gpu):
    """The 3,289,545-node layered model with 256 sources at its tolerance (1e-8): 16 sampled
    columns against the oracle's operator and against fp64 GPU potentials, the adjoint identity,
    and one Gauss-Newton iteration (10 CG steps, fields recomputed per Hessian product) giving
    the same objective in mixed precision and in fp64."""
    from simpeg_amd import maps, workloads
    from simpeg_amd.inversion import gauss_newton
    from simpeg_amd.resistivity import Simulation3DNodal

    wl = workloads.layered_ellipsoid_model(0)
    nS = wl.survey.nSrc
    cols = list(range(0, nS, nS // 16))
    orc = oracle_for(wl.mesh, wl.survey)
    orc.set_model(wl.model)
    A = orc.getA()
    q = orc.source_term()[:, cols]
    sim = Simulation3DNodal(wl.mesh, survey=wl.survey, sigmaMap=maps.ExpMap(wl.mesh), rtol=1e-8)
    f = sim.fields(wl.model)
    print("mixed solve", sim.last_solve, flush=True)
    u = _host_cols(sim, f, cols)
    res = _residuals(A, q, u)
    print("max sampled residual", res.max(), flush=True)
    assert res.max() <= 1.05e-8
    ref = Simulation3DNodal(wl.mesh, survey=wl.survey, sigmaMap=maps.ExpMap(wl.mesh),
                            precision="fp64", rtol=1e-12)
    fr = ref.fields(wl.model)
    ur = _host_cols(ref, fr, cols)
    assert max(rel(u[:, k], ur[:, k]) for k in range(len(cols))) < 1e-7
    rng = np.random.default_rng(6)
    w = rng.standard_normal(wl.survey.nD)
    v = rng.standard_normal(wl.mesh.n_cells)
    assert rel(sim.Jtvec(wl.model, w, f=f), ref.Jtvec(wl.model, w, f=fr)) < 1e-6
    sim.rtol = 1e-10
    wjv = w @ sim.Jvec(wl.model, v, f=f)
    vjtw = v @ sim.Jtvec(wl.model, w, f=f)
    assert abs(wjv - vjtw) < 1e-8 * max(abs(wjv), abs(vjtw))
    del f, fr, ref
    # one Gauss-Newton iteration from the layered background, mixed against fp64
    dobs, wd = workloads.noisy_weights(sim.dpred(wl.model))
    m0 = np.log(np.where(-wl.mesh.cell_centers[:, 2] < 100.0, 0.005,
                         np.where(-wl.mesh.cell_centers[:, 2] < 400.0, 0.02, 0.002)))
    hist = {}
    for prec in ("mixed", "fp64"):
        s = Simulation3DNodal(wl.mesh, survey=wl.survey, sigmaMap=maps.ExpMap(wl.mesh),
                              precision=prec, rtol=1e-8)
        _, h = gauss_newton(s, dobs, wd, m0, beta=1e-2, n_iter=1, cg_iter=10, recompute=True)
        hist[prec] = h
        print(prec, "phi", h.phi, flush=True)
    assert hist["mixed"].phi[-1] < hist["mixed"].phi[0]
    assert np.allclose(hist["mixed"].phi, hist["fp64"].phi, rtol=1e-6)
