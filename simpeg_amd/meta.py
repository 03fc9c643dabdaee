"""Many simulations behind one simulation interface (simpeg/meta/simulation.py:13-330).

``MetaSimulation(simulations, mappings)``: fields / dpred / Jvec are per-simulation results
concatenated, Jtvec = sum_i mapping_i'^T Jtvec_i, getJtJdiag sums the per-simulation diagonals
through the mappings (meta/simulation.py:194-330).

On the GPU, sub-simulations that share a mesh, boundary condition and solver settings form a
group: one engine holds one operator per problem (operator sets), and each group's fields,
Jvec and Jtvec are one batched solve over the sources of all its problems (every aligned group
of 8 columns follows its own problem's operator).  A 48-problem batch on five mesh sizes thus
runs five large solves instead of 48 small ones.  Groups and the remaining single simulations
run concurrently, each driven from its own host thread on its own HIP stream (the C library
releases the GIL), largest first.
"""

from __future__ import annotations

import concurrent.futures as cf
import warnings

import numpy as np
import torch

from .engine import NodalEngine
from .engine import download as _download
from .maps import ExpMap, IdentityMap, Projection, ReciprocalMap

# elementwise maps whose device evaluation has no per-instance state
_STATELESS_MAPS = (IdentityMap, ExpMap, ReciprocalMap)


def _tiles(mappings, n):
    """True when every mapping is a contiguous Projection and their ranges cover [0, n) once."""
    rs = []
    for mp in mappings:
        if not (isinstance(mp, Projection) and mp._range is not None):
            return False
        rs.append(mp._range.indices(n)[:2])
    pos = 0
    for a, b in sorted(rs):
        if a != pos:
            return False
        pos = b
    return pos == n


def sum_adjoint(mappings, model, parts):
    """sum_i mapping_i'(m)^T part_i, accumulated in place: a Projection with distinct indices
    adds its part into its own entries (no full-length temporary per problem); contiguous
    Projections that tile the model are written, not added."""
    n = np.asarray(model).size
    if _tiles(mappings, n):
        out = np.empty(n)
        for mp, part in zip(mappings, parts):
            out[mp._range] = part
        return out
    out = np.zeros(n)
    for mp, part in zip(mappings, parts):
        if isinstance(mp, Projection) and mp._range is not None:
            out[mp._range] += part
        elif isinstance(mp, Projection) and mp.unique_index:
            out[mp.index] += part
        else:
            out += mp.deriv(model).T @ part
    return out


def _group_key(sim):
    """Simulations with equal keys can share one engine (same mesh, boundary and solver
    settings); None: the simulation runs on its own engine."""
    if sim.storeJ or sim.miniaturize or sim.surface_faces is not None:
        return None
    return (tuple(np.asarray(h, dtype=float).tobytes() for h in sim._widths),
            np.asarray(sim._origin, dtype=float).tobytes(), sim._bc_type, sim.precision,
            sim.rtol, sim.inner_rtol, sim.max_outer, sim.max_iter, str(sim._device),
            sim._coarse_max, tuple(sorted(sim._gpu_params.items())), sim._deriv_includes_bc,
            sim.autotune)


class _Single:
    """One simulation on its own engine."""

    def __init__(self, i, sim):
        self.members = [i]
        self.sim = sim
        self.cost = sim._nN * sim.survey.nSrc

    def fields(self):
        return [self.sim.fields(self.sim.model)]

    def dpred(self, f):
        return [self.sim.dpred(self.sim.model, f=f[0])]

    def jvec(self, v, f):
        return [self.sim.Jvec(self.sim.model, v[0], f=f[0])]

    def jtvec(self, v, f):
        return [self.sim.Jtvec(self.sim.model, v[0], f=f[0])]


def _host_buffer(shape):
    """Host float64 storage for the private model copy: a page-locked torch tensor when a GPU is
    present (the groups upload their members' blocks of it asynchronously), else an array."""
    if torch.cuda.is_available():
        t = torch.empty(shape, dtype=torch.float64, pin_memory=True)
        # touched once here: the first write to fresh page-locked memory is ~10x slower than
        # later ones (19 against 1.9 ms for 55 MB), which would land on the next model set
        t.numpy().fill(0.0)
        return t
    return np.empty(shape)


class _SetGroup:
    """Simulations on one mesh: one engine with one operator set per simulation; the sources of
    simulation k occupy columns offs[k] .. offs[k] + counts[k] of every block, each simulation's
    block padded to a multiple of 8 columns with zero right-hand sides."""

    def __init__(self, members, sims):
        self.members = list(members)
        self.sims = list(sims)
        s0 = self.sims[0]
        self.engine = NodalEngine(
            s0._widths, s0._origin, bc_type="Neumann" if s0._bc_type == "Neumann" else "Robin",
            device=s0._device, coarse_max=s0._coarse_max, params=s0._gpu_params)
        self.counts = [s._sim_survey.nSrc for s in self.sims]
        self.offs, colset = [], []
        for k, n in enumerate(self.counts):
            self.offs.append(len(colset))
            colset += [k] * ((n + 7) // 8 * 8)
        self.colset = np.asarray(colset, dtype=np.int32)
        self.ncol = int(self.colset.size)
        self.cost = sum(s._nN * n for s, n in zip(self.sims, self.counts))
        self._q = None
        self._tuned = False
        self._versions = None  # the members' model versions the engine's operators belong to
        self._m_dev = None     # the members' models on the device (device-evaluated maps)
        self._scat = None      # P^T of all members into the group's block (one scatter table)

    def _views(self, block):
        return [block[o:o + n] for o, n in zip(self.offs, self.counts)]

    def _device_maps(self):
        fms = [s._device_map() for s in self.sims]
        return fms if all(f is not None for f in fms) else None

    def _shared_device_map(self):
        """One device map for all members, evaluated on the (members, cells) block at once: every
        member has the same elementwise map class without per-instance state (Identity, Exp,
        Reciprocal) as its sigmaMap, or every member as its rhoMap.  None otherwise."""
        kind = (type(self.sims[0].sigmaMap), type(self.sims[0].rhoMap))
        if any((type(s.sigmaMap), type(s.rhoMap)) != kind for s in self.sims):
            return None
        if not all(t in _STATELESS_MAPS or t is type(None) for t in kind):
            return None
        fms = self._device_maps()
        return None if fms is None else fms[0]

    def _set_models(self):
        """Operators of the members' current models (rebuilt only when a model changed).  With
        elementwise maps the models go to the device and sigma is evaluated there; otherwise
        sigma is evaluated on the host."""
        eng = self.engine
        versions = tuple(s._model_version for s in self.sims)
        if versions != self._versions:
            fm = self._shared_device_map()
            fms = self._device_maps() if fm is None else None
            if fm is not None or fms is not None:
                # the members' models uploaded into one device block; a model that is a view
                # of MetaSimulation's pinned copy goes without a host copy or a wait
                m_dev = torch.empty((len(self.sims), eng.nC), dtype=torch.float64,
                                    device=eng.device)
                for k, s in enumerate(self.sims):
                    with warnings.catch_warnings():
                        # the private copy is read-only; the tensor is only read
                        warnings.simplefilter("ignore", UserWarning)
                        src = torch.from_numpy(np.ascontiguousarray(s.model, dtype=float))
                    m_dev[k].copy_(src, non_blocking=src.is_pinned())
                if fm is not None:
                    sig = fm[0](m_dev)
                else:
                    sig = torch.empty_like(m_dev)
                    for k, f in enumerate(fms):
                        sig[k] = f[0](m_dev[k])
                eng.set_sigma_sets(sig)
                self._m_dev = m_dev
            else:
                eng.set_sigma_sets(np.stack([np.asarray(s.sigma, dtype=float) for s in self.sims]))
                self._m_dev = None
            self._versions = versions
        want = 1 if self.sims[0]._deriv_includes_bc else 0
        if eng.params.deriv_robin != want:
            eng.set_params(deriv_robin=want)

    def _group_scatter(self):
        """The members' P^T tables joined into one for the group's block: member k's
        destinations shifted by its first column, its data indices by its data offset."""
        if self._scat is None:
            from .engine import ScatterCSR

            eng = self.engine
            dst, rowptr, src, w = [], [], [], []
            ne = 0
            nd = 0
            nrow = 0
            for s, off in zip(self.sims, self.offs):
                sc = s._projection(eng)["scatter"]
                dst.append(sc.dst + off * eng.ld)
                rowptr.append(sc.rowptr[:-1] + ne)
                src.append(sc.src + nd)
                w.append(sc.w)
                ne += int(sc.rowptr[-1].item())
                nd += s.survey.nD
                nrow += sc.nrow
            rowptr.append(torch.tensor([ne], dtype=torch.int32, device=eng.device))
            self._scat = ScatterCSR(nrow, torch.cat(dst), torch.cat(rowptr).to(torch.int32),
                                    torch.cat(src).to(torch.int32), torch.cat(w))
        return self._scat

    def _source(self):
        if self._q is None:
            q = self.engine.block(self.ncol)
            for s, v in zip(self.sims, self._views(q)):
                s._source_block(self.engine, out=v)
            self._q = q
        return self._q

    def _solve(self, rhs):
        s0 = self.sims[0]
        u, res = self.engine.solve_sets(rhs, self.colset, self.ncol, mode=s0.precision,
                                        tol=s0.rtol, inner_tol=s0.inner_rtol,
                                        max_outer=s0.max_outer, max_iter=s0.max_iter)
        for s in self.sims:
            s.last_solve = res
            s.solve_count += 1
            s.solve_iterations += res.iters
        return u

    def fields(self):
        from .resistivity.fields import Fields3DNodal

        self._set_models()
        q = self._source()
        s0 = self.sims[0]
        if s0.autotune and not self._tuned:
            from .autotune import tune

            tune(self.engine, q, self.ncol, mode="fp64" if s0.precision == "fp64" else "mixed",
                 tol=s0.rtol)
            self._tuned = True
        u = self._solve(q)
        return [Fields3DNodal(s, v, engine=self.engine) for s, v in zip(self.sims, self._views(u))]

    def dpred(self, f):
        eng = self.engine
        return [eng.gather(s._projection(eng)["gather"], fk.u_block).cpu().numpy()
                for s, fk in zip(self.sims, f)]

    def jvec(self, v, f):
        eng = self.engine
        self._set_models()
        rhs = eng.block(self.ncol)
        for s, vk, fk, out in zip(self.sims, v, f, self._views(rhs)):
            eng.jvec_rhs(s.sigmaDeriv @ np.asarray(vk, dtype=float), fk.u_block, out.shape[0],
                         out=out)
        du = self._solve(rhs)
        return [eng.gather(s._projection(eng)["gather"], d).cpu().numpy()
                for s, d in zip(self.sims, self._views(du))]

    def jtvec(self, v, f):
        eng = self.engine
        self._set_models()
        for s, vk in zip(self.sims, v):
            if np.size(vk) != s.survey.nD:
                raise ValueError(f"v must have {s.survey.nD} entries, got {np.size(vk)}")
        # every member's data vector in one upload, P^T of all members in one scatter
        v_dev = torch.as_tensor(np.concatenate([np.asarray(vk, dtype=float).ravel() for vk in v]))
        rhs = eng.scatter(self._group_scatter(), v_dev.to(eng.device), self.ncol)
        lam = self._solve(rhs)
        G = torch.empty((len(self.sims), eng.nC), dtype=torch.float64, device=eng.device)
        for k, (fk, lk) in enumerate(zip(f, self._views(lam))):
            eng.contract(fk.u_block, lk, lk.shape[0], out=G[k])
        if self._m_dev is not None:
            # -sigmaDeriv^T g on the device (diagonal derivative), one download for all members
            fm = self._shared_device_map()
            if fm is not None:
                flat = _download(-(fm[1](self._m_dev) * G))
            else:
                fms = self._device_maps()
                flat = _download(torch.stack([-(fms[k][1](self._m_dev[k]) * G[k])
                                              for k in range(len(self.sims))]))
            return [flat[k] for k in range(len(self.sims))]
        Gh = _download(G)
        return [-s._sigma_deriv_T(Gh[k]) for k, s in enumerate(self.sims)]


class MetaSimulation:
    def __init__(self, simulations, mappings, n_streams=8, group=True):
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
        self._model_bufs = []  # the two stores the private model copy alternates between
        self._model_buf = None  # the store holding the current model
        self._streams = None
        vnD = [sim.survey.nD for sim in self.simulations]
        self._data_offsets = np.cumsum(np.r_[0, vnD])
        # work units: groups of same-mesh simulations (two or more) and single simulations
        keys = {}
        for i, sim in enumerate(self.simulations):
            k = _group_key(sim) if group else None
            keys.setdefault(k if k is not None else ("single", i), []).append(i)
        self._units = []
        for members in keys.values():
            if len(members) > 1:
                self._units.append(_SetGroup(members, [self.simulations[i] for i in members]))
            else:
                self._units.append(_Single(members[0], self.simulations[members[0]]))
        self._units.sort(key=lambda u: -u.cost)

    @property
    def nD(self):
        return int(self._data_offsets[-1])

    @property
    def model(self):
        return self._model

    @model.setter
    def model(self, value):
        prev = self._model
        new = None if value is None else np.asarray(value, dtype=float).ravel()
        if (prev is not None and new is not None and prev.shape == new.shape
                and np.array_equal(prev[:4096], new[:4096]) and np.array_equal(prev, new)):
            return  # the same model: every simulation keeps its own
        # one private, read-only copy of the full model; a simulation behind a contiguous
        # Projection keeps a view of its block of it instead of a copy of its own.  Two buffers
        # alternate (the simulations compare the new model with the views they hold of the
        # previous one), so that a new model is copied into memory already paged in.
        if new is None:
            self._model = None
            self._model_buf = None
        else:
            cur = self._model_buf
            store = next((b for b in self._model_bufs
                          if b is not cur and tuple(b.shape) == new.shape), None)
            if store is None:
                # both stores at once: page-locking memory is slow, and the second model set
                # (typically the first timed one) then finds its store ready
                keep = [b for b in self._model_bufs if b is cur and tuple(b.shape) == new.shape]
                store = _host_buffer(new.shape)
                spare = keep[:1] or [_host_buffer(new.shape)]
                self._model_bufs = spare + [store]
            # a fresh writeable view of the store for every model (a view of page-locked torch
            # memory cannot be made writeable again once it was read-only)
            buf = store.numpy() if isinstance(store, torch.Tensor) else store
            if not buf.flags.writeable:
                buf.flags.writeable = True
            np.copyto(buf, new)
            buf.flags.writeable = False
            self._model = buf
            self._model_buf = store
        for mp, sim in zip(self.mappings, self.simulations):
            if value is None:
                sim.model = None
            elif (isinstance(mp, Projection) and mp._range is not None
                    and hasattr(sim, "_assign_model")):
                sim._assign_model(self._model[mp._range], copy=False)
            else:
                sim.model = mp * self._model

    def _map(self, fn):
        """fn(unit) for every work unit, concurrently on separate streams (longest first);
        returns the per-simulation results in simulation order."""
        units = self._units
        if not torch.cuda.is_available() or self.n_streams == 1 or len(units) == 1:
            outs = [fn(u) for u in units]
        else:
            if self._streams is None:
                self._streams = [torch.cuda.Stream() for _ in range(self.n_streams)]
            outs = [None] * len(units)

            def work(slot, idx_list):
                with torch.cuda.stream(self._streams[slot]):
                    for j in idx_list:
                        outs[j] = fn(units[j])
                    torch.cuda.current_stream().synchronize()

            load = [0.0] * self.n_streams
            plan = [[] for _ in range(self.n_streams)]
            for j, u in enumerate(units):
                s = int(np.argmin(load))
                plan[s].append(j)
                load[s] += u.cost
            with cf.ThreadPoolExecutor(self.n_streams) as ex:
                futs = [ex.submit(work, s, plan[s]) for s in range(self.n_streams) if plan[s]]
                for fu in futs:
                    fu.result()
        res = [None] * len(self.simulations)
        for u, o in zip(units, outs):
            for i, r in zip(u.members, o):
                res[i] = r
        return res

    def fields(self, m):
        self.model = m
        return self._map(lambda u: u.fields())

    def dpred(self, m=None, f=None):
        if f is None:
            f = self.fields(self._model if m is None else m)
        parts = self._map(lambda u: u.dpred([f[i] for i in u.members]))
        return np.concatenate(parts)

    def Jvec(self, m, v, f=None):
        self.model = m
        if f is None:
            f = self.fields(m)
        vm = [mp.deriv(self._model) @ v for mp in self.mappings]
        parts = self._map(lambda u: u.jvec([vm[i] for i in u.members], [f[i] for i in u.members]))
        return np.concatenate(parts)

    def Jtvec(self, m, v, f=None):
        self.model = m
        if f is None:
            f = self.fields(m)
        v = np.asarray(v, dtype=float)
        vs = [v[self._data_offsets[i]:self._data_offsets[i + 1]]
              for i in range(len(self.simulations))]
        parts = self._map(lambda u: u.jtvec([vs[i] for i in u.members], [f[i] for i in u.members]))
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
