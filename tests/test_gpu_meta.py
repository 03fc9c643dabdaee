"""MetaSimulation over independent nodal problems driven concurrently on separate HIP streams
(simpeg/meta/simulation.py semantics: concatenated data, summed Jtvec)."""

import numpy as np
import pytest

from tests._helpers import oracle_for, rel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def batch():
    from simpeg_amd import workloads as W

    return W.uneven_batch(seed=1, n_problems=5, sides=(8, 12, 16), npad=2, min_src=2, max_src=6)


def _meta(batch, n_streams):
    from simpeg_amd import maps
    from simpeg_amd.meta import MetaSimulation
    from simpeg_amd.resistivity import Simulation3DNodal

    sims, mappings, offs = [], [], [0]
    for wl in batch:
        offs.append(offs[-1] + wl.mesh.n_cells)
    nP = offs[-1]
    for k, wl in enumerate(batch):
        sims.append(Simulation3DNodal(wl.mesh, survey=wl.survey, sigmaMap=maps.ExpMap(wl.mesh)))
        mappings.append(maps.Projection(nP, slice(offs[k], offs[k + 1])))
    m = np.concatenate([wl.model for wl in batch])
    return MetaSimulation(sims, mappings, n_streams=n_streams), m


def test_meta_concurrent_equals_individual(gpu, batch):
    meta4, m = _meta(batch, 4)
    meta1, _ = _meta(batch, 1)
    d4 = meta4.dpred(m)
    d1 = meta1.dpred(m)
    d_ref = np.concatenate([oracle_for(wl.mesh, wl.survey).dpred(wl.model) for wl in batch])
    assert rel(d4, d_ref) < 1e-6
    assert rel(d4, d1) < 1e-9
    rng = np.random.default_rng(0)
    v = rng.standard_normal(m.size)
    w = rng.standard_normal(d4.size)
    f4 = meta4.fields(m)
    jv = meta4.Jvec(m, v, f=f4)
    jtw = meta4.Jtvec(m, w, f=f4)
    assert abs(w @ jv - v @ jtw) < 1e-8 * abs(w @ jv)
    f1 = meta1.fields(m)
    assert rel(jtw, meta1.Jtvec(m, w, f=f1)) < 1e-9
    # Jtvec of the concatenated model = per-problem Jtvec blocks
    off = 0
    doff = 0
    for wl in batch:
        orc = oracle_for(wl.mesh, wl.survey)
        n, nd = wl.mesh.n_cells, wl.survey.nD
        assert rel(jtw[off:off + n], orc.Jtvec(wl.model, w[doff:doff + nd])) < 1e-6
        off += n
        doff += nd
