"""Many simulations behind one simulation interface (simpeg/meta/simulation.py:13-330).

``MetaSimulation(simulations, mappings)``: fields / dpred / Jvec are per-simulation results
concatenated, Jtvec = sum_i mapping_i'^T Jtvec_i, getJtJdiag sums the per-simulation diagonals
through the mappings.  On the GPU the simulations run concurrently: each one is driven from its
own host thread on its own HIP stream (the C library releases the GIL, and kernels of different
streams overlap on the device), largest problems first.  With 4 hardware queues per process,
8 streams still helped on MI355X: while one problem's host thread waits on its convergence poll
another's kernels fill the queue (48 uneven problems: fields 150 ms with 4 streams, 130 ms with
8, 129 ms with 12).
"""

from __future__ import annotations

import concurrent.futures as cf

import numpy as np
import torch

from .maps import Projection


def sum_adjoint(mappings, model, parts):
    """sum_i mapping_i'(m)^T part_i, accumulated in place: a Projection with distinct indices
    adds its part into its own entries (no full-length temporary per problem)."""
    out = np.zeros(np.asarray(model).size)
    for mp, part in zip(mappings, parts):
        if isinstance(mp, Projection) and mp.unique_index:
            out[mp.index] += part
        else:
            out += mp.deriv(model).T @ part
    return out


class MetaSimulation:
    def __init__(self, simulations, mappings, n_streams=8):
        if len(simulations) != len(mappings):
            raise ValueError("Must provide the same number of mappings and simulations.")
        ids = [id(s) for s in simulations]
        if len(set(ids)) != len(ids):
            raise ValueError("simulations must be unique")
        nP = {getattr(mp, "nP", "*") for mp in mappings}
        if len(nP) != 1:
            raise ValueError("All mappings must have the same input length")
        self.simulations = list(simulations)
        self.mappings = list(mappings)
        self.n_streams = max(1, int(n_streams))
        self._model = None
        self._streams = None
        vnD = [sim.survey.nD for sim in self.simulations]
        self._data_offsets = np.cumsum(np.r_[0, vnD])
        cost = [sim._nN * sim.survey.nSrc for sim in self.simulations]
        self._order = list(np.argsort(cost)[::-1])

    @property
    def nD(self):
        return int(self._data_offsets[-1])

    @property
    def model(self):
        return self._model

    @model.setter
    def model(self, value):
        self._model = None if value is None else np.asarray(value, dtype=float).ravel()
        for mp, sim in zip(self.mappings, self.simulations):
            sim.model = None if value is None else mp * self._model

    def _map(self, fn):
        """fn(i) for every simulation, concurrently on separate streams; results in order."""
        if not torch.cuda.is_available() or self.n_streams == 1:
            return [fn(i) for i in range(len(self.simulations))]
        if self._streams is None:
            self._streams = [torch.cuda.Stream() for _ in range(self.n_streams)]
        out = [None] * len(self.simulations)

        def work(slot, idx_list):
            with torch.cuda.stream(self._streams[slot]):
                for i in idx_list:
                    out[i] = fn(i)
                torch.cuda.current_stream().synchronize()

        # longest-processing-time-first assignment to the stream slots
        load = [0.0] * self.n_streams
        plan = [[] for _ in range(self.n_streams)]
        for i in self._order:
            s = int(np.argmin(load))
            plan[s].append(i)
            load[s] += self.simulations[i]._nN * self.simulations[i].survey.nSrc
        with cf.ThreadPoolExecutor(self.n_streams) as ex:
            futs = [ex.submit(work, s, plan[s]) for s in range(self.n_streams) if plan[s]]
            for fu in futs:
                fu.result()
        return out

    def fields(self, m):
        self.model = m
        return self._map(lambda i: self.simulations[i].fields(self.simulations[i].model))

    def dpred(self, m=None, f=None):
        if f is None:
            f = self.fields(self._model if m is None else m)
        parts = self._map(lambda i: self.simulations[i].dpred(self.simulations[i].model, f=f[i]))
        return np.concatenate(parts)

    def Jvec(self, m, v, f=None):
        self.model = m
        if f is None:
            f = self.fields(m)
        parts = self._map(lambda i: self.simulations[i].Jvec(
            self.simulations[i].model, self.mappings[i].deriv(self._model) @ v, f=f[i]))
        return np.concatenate(parts)

    def Jtvec(self, m, v, f=None):
        self.model = m
        if f is None:
            f = self.fields(m)
        v = np.asarray(v, dtype=float)
        parts = self._map(lambda i: self.simulations[i].Jtvec(
            self.simulations[i].model, v[self._data_offsets[i]:self._data_offsets[i + 1]], f=f[i]))
        return sum_adjoint(self.mappings, self._model, parts)

    def getJtJdiag(self, m, W=None, f=None):
        self.model = m
        w = np.ones(self.nD) if W is None else np.asarray(W.diagonal())
        import scipy.sparse as sp

        out = 0
        for i, sim in enumerate(self.simulations):
            wi = sp.diags(w[self._data_offsets[i]:self._data_offsets[i + 1]])
            row = sim.getJtJdiag(sim.model, W=wi, f=None if f is None else f[i])
            d = self.mappings[i].deriv(self._model)
            out = out + np.asarray((sp.diags(np.sqrt(row)) @ d).power(2).sum(axis=0)).ravel()
        return out
