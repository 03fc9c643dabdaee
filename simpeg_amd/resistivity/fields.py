"""Fields of the nodal DC simulation (electromagnetics/static/resistivity/fields.py:186-257).

The potentials stay on the GPU as a (nSrc, ld) float64 block; indexing ``f[src, name]``
copies the requested columns to the host once.  ``phiSolution``/``phi`` are the nodal
potentials; ``e = -G phi`` on edges, ``j = MeI MeSigma e``, ``charge = -eps0 G^T M_e e`` and
``charge_density = aveN2CC charge / V`` are derived on the host.
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp

EPSILON_0 = 8.8541878128e-12


class Fields3DNodal:
    knownFields = {"phiSolution": "N"}
    aliasFields = {
        "phi": ["phiSolution", "N", "_phi"],
        "j": ["phiSolution", "E", "_j"],
        "e": ["phiSolution", "E", "_e"],
        "charge_density": ["phiSolution", "CC", "_charge_density"],
        "charge": ["phiSolution", "N", "_charge"],
    }
    dtype = float

    def __init__(self, simulation, u_block=None, engine=None):
        self.simulation = simulation
        # the engine whose block holds the potentials (a MetaSimulation's shared engine, or the
        # simulation's own)
        self._eng = engine
        self.mesh = simulation.mesh
        # FieldsDC.survey: the mini survey when the simulation is miniaturised (fields.py:13-17)
        self.survey = simulation._sim_survey
        self.u_block = u_block
        self._host = None

    # ---- storage
    @property
    def phi_host(self):
        if self._host is None:
            eng = self._eng if self._eng is not None else self.simulation.engine
            self._host = eng.from_block(self.u_block, self.survey.nSrc)
        return self._host

    def _cols(self, key):
        srcs = self.survey.source_list
        if isinstance(key, slice):
            return list(range(len(srcs)))[key]
        if isinstance(key, (list, tuple)):
            return [self._src_index(s) for s in key]
        return [self._src_index(key)]

    def _src_index(self, src):
        for i, s in enumerate(self.survey.source_list):
            if s is src:
                return i
        raise KeyError("Source is not part of the survey")

    def __getitem__(self, key):
        if not isinstance(key, tuple) or len(key) != 2:
            raise KeyError("index with f[source, field_name]")
        src, name = key
        cols = self._cols(src)
        if name not in self.knownFields and name not in self.aliasFields:
            raise KeyError(f"Field {name!r} is not known")
        phi = self.phi_host[:, cols]
        sim = self.simulation
        if name in ("phiSolution", "phi"):
            out = phi
        else:
            G = self.mesh.nodal_gradient
            e = -(G @ phi)
            if name == "e":
                out = e
            elif name == "j":
                out = sp.diags(sim._edge_mass(sim.sigma) / sim._edge_mass(None)) @ e
            elif name == "charge":
                out = -EPSILON_0 * (G.T @ (sp.diags(sim._edge_mass(None)) @ e))
            else:  # charge_density
                ch = -EPSILON_0 * (G.T @ (sp.diags(sim._edge_mass(None)) @ e))
                out = (self.mesh.average_node_to_cell @ ch) / self.mesh.cell_volumes[:, None]
        # one column per requested source, also for a single source: the reference indexes
        # with a list of source indices (simpeg/fields.py:322-341), giving (n, 1)
        return out

    def __setitem__(self, key, value):
        src, name = key
        if name != "phiSolution":
            raise KeyError("only phiSolution can be assigned")
        cols = self._cols(src)
        host = self.phi_host.copy()
        value = np.asarray(value, dtype=float)
        host[:, cols] = value.reshape(host.shape[0], -1)
        eng = self._eng if self._eng is not None else self.simulation.engine
        self.u_block = eng.to_block(host)
        self._host = host
