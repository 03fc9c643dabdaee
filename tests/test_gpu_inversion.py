"""Gauss-Newton harness (batched Jvec/Jtvec Hessian products) on the GPU against the same
iteration driven by the FP64 CPU oracle."""

import numpy as np
import pytest

from tests._helpers import oracle_for, rel, tiny_problem

pytestmark = pytest.mark.gpu


class _OracleSim:
    """Adapter giving the oracle the fields/dpred/Jvec/Jtvec call shape of the simulation."""

    def __init__(self, orc):
        self.orc = orc

    def fields(self, m):
        return (np.array(m, copy=True), self.orc.fields(m))

    def dpred(self, m, f=None):
        f = self.fields(m) if f is None else f
        return self.orc.dpred(f[0], u=f[1])

    def Jvec(self, m, v, f=None):
        f = self.fields(m) if f is None else f
        return self.orc.Jvec(f[0], v, u=f[1])

    def Jtvec(self, m, v, f=None):
        f = self.fields(m) if f is None else f
        return self.orc.Jtvec(f[0], v, u=f[1])


@pytest.fixture(scope="module")
def setup():
    from simpeg_amd import workloads

    mesh, survey, m_true = tiny_problem()
    orc = oracle_for(mesh, survey)
    dclean = orc.dpred(m_true)
    dobs, wd = workloads.noisy_weights(dclean)
    m0 = np.full(mesh.n_cells, np.log(0.01))
    return mesh, survey, orc, dobs, wd, m0


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_gauss_newton_matches_oracle_run(gpu, setup, precision):
    from simpeg_amd import maps
    from simpeg_amd.inversion import gauss_newton
    from simpeg_amd.resistivity import Simulation3DNodal

    mesh, survey, orc, dobs, wd, m0 = setup
    sim = Simulation3DNodal(mesh, survey=survey, sigmaMap=maps.ExpMap(mesh), precision=precision)
    m_gpu, h_gpu = gauss_newton(sim, dobs, wd, m0, beta=1e-1, n_iter=3, cg_iter=10)
    m_cpu, h_cpu = gauss_newton(_OracleSim(orc), dobs, wd, m0, beta=1e-1, n_iter=3, cg_iter=10)
    assert h_gpu.phi[-1] < 0.5 * h_gpu.phi[0]
    assert np.allclose(h_gpu.phi, h_cpu.phi, rtol=1e-5)
    assert rel(m_gpu - m0, m_cpu - m0) < 1e-4


def test_gauss_newton_recompute_matches_stored(gpu, setup):
    from simpeg_amd import maps
    from simpeg_amd.inversion import gauss_newton
    from simpeg_amd.resistivity import Simulation3DNodal

    mesh, survey, orc, dobs, wd, m0 = setup
    sim = Simulation3DNodal(mesh, survey=survey, sigmaMap=maps.ExpMap(mesh))
    m1, h1 = gauss_newton(sim, dobs, wd, m0, beta=1e-1, n_iter=2, cg_iter=5)
    m2, h2 = gauss_newton(sim, dobs, wd, m0, beta=1e-1, n_iter=2, cg_iter=5, recompute=True)
    assert rel(m2 - m0, m1 - m0) < 1e-8
