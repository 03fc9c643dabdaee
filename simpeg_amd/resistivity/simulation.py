"""GPU ``Simulation3DNodal``: drop-in for
``simpeg.electromagnetics.static.resistivity.Simulation3DNodal``.

Reference (paths relative to the simpeg repository):
  BaseDCSimulation          electromagnetics/static/resistivity/simulation.py:20-318
  Simulation3DNodal         electromagnetics/static/resistivity/simulation.py:483-670
  BaseSimulation.dpred      simulation.py:161-196; residual 336-363; make_synthetic_data 365-433
  model / sigma properties  props.py:117-225, 408-477

The reference factorises the assembled A = G^T M_e(sigma) G + diag(AvgBC sigma) and runs one
triangular solve pair per right-hand side.  Here every solve is a batched (all sources at once)
multigrid-preconditioned CG on the GPU with a matrix-free stencil (libsimpeg_dc):

* ``fields(m)``: one batched solve over the sources (fp32 PCG inside fp64 refinement by default,
  or plain fp64 PCG with ``precision="fp64"``).
* ``Jvec``: right-hand sides -A(sigmaDeriv v) u_s built by the stencil kernel, one batched solve.
* ``Jtvec``: P^T v summed per source (A is symmetric, so the reference's one adjoint solve per
  (source, receiver) pair collapses to one per source), one batched solve, then the
  edge-gradient contraction kernel sums d(u_s^T A lam_s)/d sigma over sources.
* ``getJ``: one adjoint column per datum, per-column contraction.

Results are host numpy arrays, as in the reference.
"""

from __future__ import annotations

import warnings

import numpy as np
import scipy.sparse as sp
import torch

from .. import maps as _maps
from ..data import Data, SyntheticData
from ..engine import NodalEngine, coo_to_gather, coo_to_scatter, download
from ..mesh import edge_mass, mesh_widths, node_interpolation, robin_diagonal
from .fields import Fields3DNodal
from .survey import Survey


class Simulation3DNodal:
    """3D nodal DC problem solved on an AMD Instinct GPU.

    Parameters follow the reference (``mesh, survey, bc_type, sigma, sigmaMap, rho, rhoMap,
    solver, solver_opts, storeJ, miniaturize, surface_faces, sensitivity_path, counter,
    verbose``).  ``solver``/``solver_opts`` are accepted for signature compatibility; the GPU
    solver replaces the direct factorisation they configure.  GPU-specific keywords:

    precision : {"mixed", "mixed_ir", "fp64"}
        "mixed": fp32 multigrid PCG with fp64 reliable updates (true residual and solution
        accumulation in fp64; default); "mixed_ir": fp32 PCG restarted inside classic fp64
        iterative refinement; "fp64": plain fp64 multigrid PCG.
    rtol : float
        per-source target of |q - A u| / |q| (fp64 residual).
    inner_rtol, max_outer, max_iter : refinement / PCG limits.
    gpu_params : dict of multigrid parameters (nu_pre, nu_post, cheb_ratio, ...).
    autotune : bool
        time the candidate multigrid/launch settings on the first ``fields`` call and keep the
        fastest (cached per mesh shape in ~/.cache/simpeg_amd/autotune.json).
    exact_bc_deriv : bool
        with ``rhoMap`` and a Robin boundary, include the boundary term d(AvgBC sigma)/dm in
        Jvec / Jtvec / getJ.  The reference leaves it out for resistivity models
        (resistivity/simulation.py:566-570), and so does this class by default.
    """

    _solutionType = "phiSolution"
    _formulation = "EB"
    fieldsPair = Fields3DNodal

    def __init__(self, mesh, survey=None, bc_type="Robin", sigma=None, sigmaMap=None, rho=None,
                 rhoMap=None, solver=None, solver_opts=None, storeJ=False, miniaturize=False,
                 surface_faces=None, sensitivity_path=None, counter=None, verbose=False,
                 model=None, *, precision="mixed", rtol=1e-10, inner_rtol=1e-5, max_outer=10,
                 max_iter=1000, device=None, gpu_params=None, coarse_max=0, jchunk=512,
                 autotune=False, exact_bc_deriv=False):
        self.mesh = mesh
        self._widths, self._origin = mesh_widths(mesh)
        self._nC = int(np.prod([len(h) for h in self._widths]))
        self._nN = int(np.prod([len(h) + 1 for h in self._widths]))
        self._survey = None
        self.survey = survey
        if bc_type not in ("Neumann", "Robin", "Mixed"):
            if bc_type == "Dirichlet":
                raise ValueError("Dirichlet conditions are not supported in the Nodal formulation")
            raise ValueError(f"bc_type must be 'Neumann', 'Robin' or 'Mixed', got {bc_type!r}")
        self._bc_type = bc_type
        if not isinstance(storeJ, bool):
            raise TypeError("storeJ must be a bool")
        self.storeJ = storeJ
        if not isinstance(miniaturize, bool):
            raise TypeError("miniaturize must be a bool")
        self.miniaturize = miniaturize
        self._mini = None
        self.surface_faces = surface_faces
        self.solver = solver
        self.solver_opts = {} if solver_opts is None else dict(solver_opts)
        self.sensitivity_path = sensitivity_path
        self.counter = counter
        self.verbose = verbose
        if sigma is not None and rho is not None:
            raise ValueError("give sigma or rho, not both")
        if sigmaMap is not None and rhoMap is not None:
            raise ValueError("give sigmaMap or rhoMap, not both")
        self.sigmaMap = sigmaMap
        self.rhoMap = rhoMap
        self._sigma_fixed = None if sigma is None else np.asarray(sigma, dtype=float).ravel()
        self._rho_fixed = None if rho is None else np.asarray(rho, dtype=float).ravel()
        if precision not in ("mixed", "fp64", "mixed_ir"):
            raise ValueError("precision must be 'mixed', 'mixed_ir' or 'fp64'")
        self.precision = precision
        self.rtol = float(rtol)
        self.inner_rtol = float(inner_rtol)
        self.max_outer = int(max_outer)
        self.max_iter = int(max_iter)
        self.jchunk = int(jchunk)
        self._device = device
        # batched solves issued and their PCG iterations (instrumentation)
        self.solve_count = 0
        self.solve_iterations = 0
        self._gpu_params = dict(gpu_params or {})
        self._coarse_max = int(coarse_max)
        self._engine = None
        self.autotune = bool(autotune)
        self.exact_bc_deriv = bool(exact_bc_deriv)
        self._tuned = False
        self._model = None
        self._model_version = 0  # counts model changes (operators built elsewhere check it)
        self._sigma_on_gpu = None
        self._m_dev = None  # the model on the device when sigma was evaluated there
        self._Jmatrix = None
        self._gtgdiag = None
        self._Av = None
        self._sd_cache = None
        self._sd_diag = None
        self._proj = None
        self._proj_other = {}
        self._q = None
        self._q_block = None
        self.last_solve = None
        if miniaturize:
            from .utils import _mini_pole_pole

            self._mini_T, self._mini = _mini_pole_pole(self.survey)
        if model is not None:
            self.model = model

    # ------------------------------------------------------------------ properties
    @property
    def survey(self):
        if self._survey is None:
            raise AttributeError("Simulation must have a survey")
        return self._survey

    @survey.setter
    def survey(self, value):
        if value is not None and not isinstance(value, Survey) and not hasattr(value, "source_list"):
            raise TypeError("survey must be a DC Survey")
        self._survey = value
        self._proj = None
        self._proj_other = {}
        self._q = None
        self._q_block = None
        self._Jmatrix = None
        self._gtgdiag = None

    @property
    def bc_type(self):
        return self._bc_type

    @property
    def _sim_survey(self):
        """The survey whose sources are solved (the pole-pole mini survey when miniaturised)."""
        return self._mini if self._mini is not None else self.survey

    def _mini_data(self, d):
        if self._mini is None:
            return d
        from .utils import mini_survey_data

        return mini_survey_data(self._mini_T, d)

    def _mini_dataT(self, v):
        if self._mini is None:
            return v
        from .utils import mini_survey_dataT

        return mini_survey_dataT(self._mini_T, v)

    @property
    def _active_map(self):
        return self.sigmaMap if self.sigmaMap is not None else self.rhoMap

    @property
    def model(self):
        return self._model

    @model.setter
    def model(self, value):
        self._assign_model(value, copy=True)

    def _assign_model(self, value, copy):
        """The model setter; copy=False keeps ``value`` itself (MetaSimulation hands over
        read-only views of its own private copy of the full model)."""
        if value is None:
            self._model = None
            return
        value = np.asarray(value, dtype=float).ravel()
        mp = self._active_map
        if mp is not None:
            nP = mp.nP
            if nP != "*" and value.size != nP:
                raise ValueError(
                    f"Model is the wrong size: {value.size}, expected {nP} parameters")
        prev = self._model
        changed = prev is None or prev.shape != value.shape
        if not changed:
            # a new model usually differs in its first entries already: the prefix settles it
            # without a full pass (same outcome as np.allclose over all entries)
            k = min(prev.size, 4096)
            changed = not np.allclose(prev[:k], value[:k]) or not (
                np.array_equal(prev, value) or np.allclose(prev, value))
        if changed:
            self._sigma_on_gpu = None
            self._Jmatrix = None
            self._gtgdiag = None
            self._model_version += 1
        self._model = value.copy() if copy else value

    @property
    def sigma(self):
        if self.sigmaMap is not None:
            if self._model is None:
                raise AttributeError("a model is needed to evaluate sigmaMap")
            return np.asarray(self.sigmaMap * self._model, dtype=float).ravel()
        if self.rhoMap is not None:
            if self._model is None:
                raise AttributeError("a model is needed to evaluate rhoMap")
            return 1.0 / np.asarray(self.rhoMap * self._model, dtype=float).ravel()
        if self._sigma_fixed is not None:
            return self._sigma_fixed
        if self._rho_fixed is not None:
            return 1.0 / self._rho_fixed
        raise AttributeError("sigma is not set: give sigma, rho, sigmaMap or rhoMap")

    @property
    def rho(self):
        return 1.0 / self.sigma

    @property
    def sigmaDeriv(self):
        # kept for the current model array and maps (every model assignment stores a new array)
        key = (self._model, self.sigmaMap, self.rhoMap)
        if self._sd_cache is not None and all(a is b for a, b in zip(self._sd_cache[0], key)):
            return self._sd_cache[1]
        if self.sigmaMap is not None:
            sd = sp.csr_matrix(self.sigmaMap.deriv(self._model))
        elif self.rhoMap is not None:
            s = self.sigma
            sd = sp.csr_matrix(_maps.sdiag(-s * s) @ self.rhoMap.deriv(self._model))
        else:
            raise AttributeError("sigmaDeriv needs sigmaMap or rhoMap")
        self._sd_cache = (key, sd)
        return sd

    def _sigma_deriv_T(self, g):
        """sigmaDeriv^T g; a diagonal derivative (Exp, Identity, Reciprocal maps) as a product of
        vectors instead of a transposed sparse product."""
        sd = self.sigmaDeriv
        if self._sd_diag is None or self._sd_diag[0] is not sd:  # checked once per matrix
            n = sd.shape[0]
            diag = sd.shape[1] == n and sd.nnz == n and np.array_equal(
                sd.indptr, np.arange(n + 1)) and np.array_equal(sd.indices, np.arange(n))
            self._sd_diag = (sd, diag)
        if self._sd_diag[1]:
            return sd.data * np.asarray(g, dtype=float).ravel()
        return sd.T @ g

    # ------------------------------------------------------------------ GPU state
    @property
    def engine(self):
        if self._engine is None:
            self._engine = NodalEngine(
                self._widths, self._origin,
                bc_type="Neumann" if self._bc_type == "Neumann" else "Robin",
                surface_faces=self.surface_faces, device=self._device,
                coarse_max=self._coarse_max, params=self._gpu_params)
        return self._engine

    def _ensure_sigma(self):
        if self._sigma_on_gpu is None or self.sigmaMap is None and self.rhoMap is None:
            self._m_dev = None
            fm = self._device_map() if self._model is not None else None
            if fm is not None and self._model.size == self._nC:
                # an elementwise map: the model goes to the device once and sigma (and later
                # its derivative) is evaluated there
                with warnings.catch_warnings():
                    # a model held as a read-only view (MetaSimulation's copy) is only read
                    warnings.simplefilter("ignore", UserWarning)
                    m_t = torch.from_numpy(np.ascontiguousarray(self._model, dtype=float))
                m_t = m_t.to(self.engine.device)
                self.engine.set_sigma(fm[0](m_t).contiguous())
                self._m_dev = m_t
                self._sigma_on_gpu = True
                return
            sig = self.sigma
            if sig.size != self._nC:
                # the reference raises this for Robin (resistivity/simulation.py:546-551); its
                # Neumann path accepts tensors, which the GPU stencil does not implement
                if self._bc_type != "Neumann":
                    raise NotImplementedError(
                        "Anisotropic conductivity is not supported for Robin boundary "
                        "conditions, please use 'Neumann'.")
                raise NotImplementedError(
                    "anisotropic conductivity is not supported by the GPU nodal simulation")
            self.engine.set_sigma(sig)
            self._sigma_on_gpu = True

    @property
    def _deriv_includes_bc(self):
        """Whether dA/dm carries the Robin term d(AvgBC sigma)/dm.  The reference's getADeriv adds
        it only when sigmaMap is set (resistivity/simulation.py:566-570): with a rhoMap model it is
        left out, unless ``exact_bc_deriv`` asks for the exact derivative."""
        return self.sigmaMap is not None or self.exact_bc_deriv

    def _sync_deriv_bc(self):
        eng = self.engine
        want = 1 if self._deriv_includes_bc else 0
        if eng.params.deriv_robin != want:
            eng.set_params(deriv_robin=want)

    def _solve(self, rhs, ncol):
        u, res = self.engine.solve(rhs, ncol, mode=self.precision, tol=self.rtol,
                                   inner_tol=self.inner_rtol, max_outer=self.max_outer,
                                   max_iter=self.max_iter)
        self.last_solve = res
        self.solve_count += 1
        self.solve_iterations += res.iters
        return u

    # ------------------------------------------------------------------ survey -> sparse maps
    def _projection(self, eng=None):
        """Receiver projections (all data) and source right-hand sides as sparse COO.  ``eng``:
        build the device tables for another engine on the same mesh (MetaSimulation's shared
        engines); they are cached per engine."""
        if eng is not None and eng is not self._engine:
            key = id(eng)
            if key not in self._proj_other:
                self._proj_other[key] = self._build_projection(eng)
            return self._proj_other[key]
        if self._proj is not None:
            return self._proj
        self._proj = self._build_projection(self.engine)
        return self._proj

    def _build_projection(self, eng):
        ld = eng.ld
        rows, flat, w, nodes, dsrc = [], [], [], [], []
        i0 = 0
        for s, src in enumerate(self._sim_survey.source_list):
            for rx in src.receiver_list:
                nd = rx.nD
                factor = np.ones(nd)
                if rx.data_type == "apparent_resistivity":
                    try:
                        factor = 1.0 / np.asarray(rx.geometric_factor[src], dtype=float)
                    except KeyError:
                        raise KeyError(
                            "Receiver geometric factor has not been set, please execute "
                            "survey.set_geometric_factor()")
                for locs, sign in rx._electrode_sets():
                    idx, wt = node_interpolation(self._widths, self._origin, locs)
                    rows.append(np.repeat(i0 + np.arange(nd), 8))
                    nodes.append(idx.ravel())
                    flat.append(s * ld + idx.ravel())
                    w.append((sign * factor[:, None] * wt).ravel())
                dsrc.append(np.full(nd, s, dtype=np.int64))
                i0 += nd
        nD = i0
        rows = np.concatenate(rows) if rows else np.zeros(0, np.int64)
        flat = np.concatenate(flat) if flat else np.zeros(0, np.int64)
        nodes = np.concatenate(nodes) if nodes else np.zeros(0, np.int64)
        w = np.concatenate(w) if w else np.zeros(0)
        keep = w != 0.0
        rows, flat, nodes, w = rows[keep], flat[keep], nodes[keep], w[keep]
        data_src = np.concatenate(dsrc) if dsrc else np.zeros(0, np.int64)
        rp, idx, ww = coo_to_gather(rows, flat, w, nD)
        gather = eng.gather_csr(rp, idx, ww)
        dst, srp, ssrc, sw = coo_to_scatter(flat, rows, w)
        scatter = eng.scatter_csr(dst, srp, ssrc, sw)
        return dict(nD=nD, rows=rows, nodes=nodes, w=w, data_src=data_src, gather=gather,
                    scatter=scatter)

    def getSourceTerm(self):
        """q (n_nodes, nSrc), Fortran ordered (resistivity/simulation.py:265-289)."""
        if self._q is None:
            srcs = self._sim_survey.source_list
            q = np.zeros((self._nN, len(srcs)), order="F")
            for i, src in enumerate(srcs):
                q[:, i] = src.eval(self)
            self._q = q
        return self._q

    getRHS = getSourceTerm

    def _source_block(self, eng=None, out=None):
        """q as a device block (nSrc, ld).  ``eng``/``out``: write it into ``out`` (a block of
        another engine on the same mesh) instead of caching it on this simulation."""
        if eng is not None:
            return self._build_source_block(eng, out)
        if self._q_block is None:
            self._q_block = self._build_source_block(self.engine, None)
        return self._q_block

    def _build_source_block(self, eng, out):
        dest, val = [], []
        for i, src in enumerate(self._sim_survey.source_list):
            locs, cur = src.electrodes()
            idx, wt = node_interpolation(self._widths, self._origin, locs)
            dest.append(i * eng.ld + idx.ravel())
            val.append((cur[:, None] * wt).ravel())
        dest = np.concatenate(dest)
        val = np.concatenate(val)
        keep = val != 0.0
        dest, val = dest[keep], val[keep]
        d, rp, sidx, sw = coo_to_scatter(dest, np.zeros(dest.size, np.int64), val)
        sc = eng.scatter_csr(d, rp, sidx, sw)
        return eng.scatter(sc, np.ones(1), self._sim_survey.nSrc, out=out)

    # ------------------------------------------------------------------ public API
    def fields(self, m=None, calcJ=True):
        """Nodal potentials of every source (``BaseDCSimulation.fields``, simulation.py:100-113)."""
        if m is not None:
            self.model = m
        self._ensure_sigma()
        q = self._source_block()
        if self.autotune and not self._tuned:
            from ..autotune import tune

            tune(self.engine, q, self._sim_survey.nSrc,
                 mode="fp64" if self.precision == "fp64" else "mixed", tol=self.rtol)
            self._tuned = True
        u = self._solve(q, self._sim_survey.nSrc)
        return Fields3DNodal(self, u)

    def dpred(self, m=None, f=None):
        """Predicted data (``BaseSimulation.dpred``, simulation.py:161-196)."""
        if self._survey is None:
            raise AttributeError(
                "The survey has not yet been set and is required to compute "
                "data. Please set the survey for the simulation: "
                "simulation.survey = survey")
        if f is None:
            if m is None:
                m = self.model
            f = self.fields(m)
        pr = self._projection()
        return self._mini_data(self.engine.gather(pr["gather"], f.u_block).cpu().numpy())

    def residual(self, m, dobs, f=None):
        return self.dpred(m, f=f) - np.asarray(dobs, dtype=float).ravel()

    def Jvec(self, m, v, f=None):
        """J v (``BaseDCSimulation.Jvec``, simulation.py:156-187)."""
        if f is None:
            f = self.fields(m)
        self.model = m
        if self.storeJ:
            return self.getJ(m, f=f) @ np.asarray(v, dtype=float)
        self._ensure_sigma()
        self._sync_deriv_bc()
        dsig = self.sigmaDeriv @ np.asarray(v, dtype=float).ravel()
        eng = self.engine
        nS = self._sim_survey.nSrc
        rhs = eng.jvec_rhs(dsig, f.u_block, nS)
        du = self._solve(rhs, nS)
        return self._mini_data(eng.gather(self._projection()["gather"], du).cpu().numpy())

    def Jtvec(self, m, v, f=None):
        """J^T v (``BaseDCSimulation.Jtvec``/``_Jtvec``, simulation.py:189-263)."""
        if f is None:
            f = self.fields(m)
        self.model = m
        if isinstance(v, Data):
            v = v.dobs
        v = np.asarray(v, dtype=float).ravel()
        if self.storeJ:
            return np.asarray(self.getJ(m, f=f).T @ v)
        self._ensure_sigma()
        self._sync_deriv_bc()
        eng = self.engine
        nS = self._sim_survey.nSrc
        pr = self._projection()
        if v.size != self.survey.nD:
            raise ValueError(f"v must have {self.survey.nD} entries, got {v.size}")
        v = self._mini_dataT(v)
        rhs = eng.scatter(pr["scatter"], v, nS)
        lam = self._solve(rhs, nS)
        g = eng.contract(f.u_block, lam, nS)
        if self._m_dev is not None:
            # -dsigma/dm^T g with the diagonal derivative evaluated on the device model
            return download(-(self._device_map()[1](self._m_dev) * g))
        return -self._sigma_deriv_T(download(g))

    def getJ(self, m, f=None):
        """Dense sensitivity (nD, nP) (``getJ``, simulation.py:115-121), one adjoint column per datum."""
        self.model = m
        if self._Jmatrix is None:
            if f is None:
                f = self.fields(m)
            self._ensure_sigma()
            self._sync_deriv_bc()
            eng = self.engine
            pr = self._projection()
            nD = pr["nD"]
            sd = self.sigmaDeriv
            J = np.empty((nD, sd.shape[1]))
            rows, nodes, w = pr["rows"], pr["nodes"], pr["w"]
            for c0 in range(0, nD, self.jchunk):
                c1 = min(nD, c0 + self.jchunk)
                k = c1 - c0
                sel = (rows >= c0) & (rows < c1)
                dst = (rows[sel] - c0) * eng.ld + nodes[sel]
                d, rp, sidx, sw = coo_to_scatter(dst, rows[sel] - c0, w[sel])
                rhs = eng.scatter(eng.scatter_csr(d, rp, sidx, sw), np.ones(k), k)
                lam = self._solve(rhs, k)
                G = eng.contract(f.u_block, lam, k, ucol=pr["data_src"][c0:c1], per_column=True)
                J[c0:c1] = -(sd.T @ G.cpu().numpy().T).T
            self._Jmatrix = self._mini_data(J)
        return self._Jmatrix

    def getJtJdiag(self, m, W=None, f=None):
        """diag(J^T W^T W J) (``getJtJdiag``, simulation.py:137-154).

        With a stored J (storeJ or an earlier getJ) it is formed from J like the reference.
        Otherwise, for a diagonal map derivative, it is accumulated on the GPU over chunks of
        adjoint columns (one per datum) without materialising J."""
        if self._gtgdiag is None:
            self.model = m
            sd = self.sigmaDeriv
            diag_map = sd.shape[0] == sd.shape[1] and sd.nnz <= sd.shape[0] and np.all(
                sd.tocoo().row == sd.tocoo().col)
            if self._Jmatrix is not None or self.storeJ or not diag_map or self._mini is not None:
                J = self.getJ(m, f=f)
                w = np.ones(J.shape[0]) if W is None else np.asarray(W.diagonal()) ** 2
                self._gtgdiag = (w[:, None] * J * J).sum(axis=0)
                return self._gtgdiag
            if f is None:
                f = self.fields(m)
            self._ensure_sigma()
            self._sync_deriv_bc()
            eng = self.engine
            pr = self._projection()
            nD = pr["nD"]
            wsq = np.ones(nD) if W is None else np.asarray(W.diagonal(), dtype=float) ** 2
            acc = torch.zeros(self._nC, dtype=torch.float64, device=eng.device)
            rows, nodes, w = pr["rows"], pr["nodes"], pr["w"]
            for c0 in range(0, nD, self.jchunk):
                c1 = min(nD, c0 + self.jchunk)
                k = c1 - c0
                sel = (rows >= c0) & (rows < c1)
                dst = (rows[sel] - c0) * eng.ld + nodes[sel]
                d, rp, sidx, sw = coo_to_scatter(dst, rows[sel] - c0, w[sel])
                rhs = eng.scatter(eng.scatter_csr(d, rp, sidx, sw), np.ones(k), k)
                lam = self._solve(rhs, k)
                eng.jtj_diag_accumulate(f.u_block, lam, k, pr["data_src"][c0:c1], wsq[c0:c1], acc)
            dsig = sd.diagonal()
            self._gtgdiag = dsig * dsig * acc.cpu().numpy()
        return self._gtgdiag

    def make_synthetic_data(self, m, relative_error=0.05, noise_floor=0.0, f=None,
                            add_noise=False, random_seed=None, **kwargs):
        if kwargs.pop("std", None) is not None:
            raise TypeError("The std parameter has been removed. Please use relative_error.")
        if f is None:
            f = self.fields(m)
        dclean = self.dpred(m, f=f)
        if add_noise:
            rng = np.random.default_rng(seed=random_seed)
            std = np.sqrt((relative_error * np.abs(dclean)) ** 2 + noise_floor ** 2)
            dobs = dclean + rng.normal(loc=0, scale=std, size=dclean.shape)
        else:
            dobs = dclean
        return SyntheticData(survey=self.survey, dobs=dobs, dclean=dclean,
                             relative_error=relative_error, noise_floor=noise_floor)

    # ------------------------------------------------------------------ device-resident path
    def _device_map(self):
        """(sigma(m), diagonal of dsigma/dm) as functions of a model held in a device tensor, or
        None when the map has no device evaluation (it is not elementwise) or the survey is
        miniaturised (its recombination is a host sparse matrix)."""
        if self._mini is not None:
            return None
        if self.sigmaMap is not None:
            mp, recip = self.sigmaMap, False
        elif self.rhoMap is not None:
            mp, recip = self.rhoMap, True
        else:
            return None
        tf, td = getattr(mp, "torch_transform", None), getattr(mp, "torch_deriv_diag", None)
        if tf is None or td is None or mp.nP not in ("*", self._nC):
            return None
        if not recip:
            return tf, td

        def sig(m):
            return 1.0 / tf(m)

        def dsig(m):
            s = 1.0 / tf(m)
            return -s * s * td(m)

        return sig, dsig

    def _device_model(self, m_t):
        fm = self._device_map()
        if fm is None:
            raise ValueError("the model map has no device evaluation (use the host API)")
        m_t = m_t.detach().to(self.engine.device, torch.float64).contiguous()
        if m_t.numel() != self._nC:
            raise ValueError(f"Model is the wrong size: {m_t.numel()}, expected {self._nC}")
        return fm, m_t

    def dpred_device(self, m_t):
        """Predicted data of a model held in a device tensor, as a device tensor, without host
        copies of the model or the data.  Returns (d, state); state = (sigma, potentials) is what
        ``jtvec_device`` reuses."""
        fm, m_t = self._device_model(m_t)
        eng = self.engine
        sig = fm[0](m_t).contiguous()
        eng.set_sigma(sig)
        # the engine's operator no longer belongs to the host-side model
        self._sigma_on_gpu = None
        u = self._solve(self._source_block(), self._sim_survey.nSrc)
        d = eng.gather(self._projection()["gather"], u)
        return d, (sig, u)

    def jtvec_device(self, m_t, v_t, state=None):
        """J^T v for a device model and a device data vector, as a device tensor.  ``state`` from
        ``dpred_device`` at the same model reuses its potentials; None solves them again."""
        fm, m_t = self._device_model(m_t)
        eng = self.engine
        if state is None:
            _, state = self.dpred_device(m_t)
        sig, u = state
        if eng._sigma is not sig:
            eng.set_sigma(sig)
            self._sigma_on_gpu = None
        self._sync_deriv_bc()
        nS = self._sim_survey.nSrc
        v_t = v_t.detach().to(eng.device, torch.float64).contiguous()
        if v_t.numel() != self.survey.nD:
            raise ValueError(f"v must have {self.survey.nD} entries, got {v_t.numel()}")
        lam = self._solve(eng.scatter(self._projection()["scatter"], v_t, nS), nS)
        g = eng.contract(u, lam, nS)
        return -(fm[1](m_t) * g)

    # ------------------------------------------------------------------ operators on the host
    def _edge_mass(self, sigma):
        return edge_mass(self._widths, sigma)

    @property
    def MeSigma(self):
        return sp.diags(self._edge_mass(self.sigma), format="csr")

    def getA(self, resistivity=None):
        """Assembled A (CSR), for inspection; the GPU path never assembles it."""
        sig = self.sigma if resistivity is None else 1.0 / np.asarray(resistivity, dtype=float)
        G = self.mesh.nodal_gradient if hasattr(self.mesh, "nodal_gradient") else None
        if G is None:
            raise AttributeError("the mesh provides no nodal_gradient")
        A = (G.T @ sp.diags(self._edge_mass(sig)) @ G).tocsr()
        if self._bc_type == "Neumann":
            A = A.tolil()
            A[0, :] = 0.0
            A[0, 0] = 1.0
            return A.tocsr()
        from ..mesh import robin_face_weights
        beta = robin_face_weights(self._widths, self._origin, self.surface_faces)
        return (A + sp.diags(robin_diagonal(self._widths, beta, sig))).tocsr()

    def getADeriv(self, u, v, adjoint=False):
        """dA/dm products (resistivity/simulation.py:557-571 with _inner_mat_mul_op,
        base/pde_simulation.py:15-83), evaluated on the GPU.

        Forward: u (nN,) or (nN, k), v (nP,) -> (nN,) or (nN, k).  Adjoint, with u and v squeezed
        as the reference does: u (nN,) and v (nN,) -> (nP,); u (nN,) and v (nN, k) -> (nP, k), one
        column per column of v; u (nN, k) and v (nN, k) -> (nP,), summed over the k pairs."""
        self._ensure_sigma()
        self._sync_deriv_bc()
        eng = self.engine
        u = np.squeeze(np.asarray(u, dtype=float))
        if u.ndim not in (1, 2) or u.shape[0] != self._nN:
            raise ValueError(f"u must have {self._nN} rows, got shape {np.shape(u)}")
        ub = eng.to_block(u)
        if not adjoint:
            dsig = self.sigmaDeriv @ np.asarray(v, dtype=float).ravel()
            out = -eng.from_block(eng.jvec_rhs(dsig, ub, ub.shape[0]))
            return out[:, 0] if u.ndim == 1 else out
        v = np.squeeze(np.asarray(v, dtype=float))
        if v.ndim not in (1, 2) or v.shape[0] != self._nN:
            raise ValueError(f"v must have {self._nN} rows, got shape {np.shape(v)}")
        lb = eng.to_block(v)
        k = lb.shape[0]
        if u.ndim == 1 and v.ndim == 2:
            # every column of v is paired with the single u: per-column products
            G = eng.contract(ub, lb, k, ucol=np.zeros(k, dtype=np.int32), per_column=True)
            return np.asarray(self.sigmaDeriv.T @ G.cpu().numpy().T)
        if u.ndim == 2 and (v.ndim != 2 or v.shape[1] != u.shape[1]):
            raise ValueError(
                f"u with {u.shape[1]} columns needs v with as many columns, got shape {v.shape}")
        g = eng.contract(ub, lb, k).cpu().numpy()
        return self.sigmaDeriv.T @ g

    def getRHS(self):
        """Right-hand sides q (nN, nSrc) (``getRHS``, simulation.py:646-653)."""
        return self.getSourceTerm()

    def MeSigmaDeriv(self, u, v=None, adjoint=False):
        """Derivative of M_e(sigma) u with respect to the model (base/pde_simulation.py:296-321,
        applied as in _inner_mat_mul_op, 15-83): the matrix diag(u) A_v sigmaDeriv, or its product
        with v (forward: u * (A_v sigmaDeriv v); adjoint: sigmaDeriv^T A_v^T (u * v))."""
        from ..mesh import edge_cell_weights

        if self._Av is None:
            self._Av = edge_cell_weights(self._widths)
        u = np.asarray(u, dtype=float)
        if v is None:
            return _maps.sdiag(u) @ self._Av @ self.sigmaDeriv
        v = np.asarray(v, dtype=float)
        if adjoint:
            return self.sigmaDeriv.T @ (self._Av.T @ (u * v))
        return u * (self._Av @ (self.sigmaDeriv @ v))

    def getRHSDeriv(self, source, v, adjoint=False):
        return 0.0
