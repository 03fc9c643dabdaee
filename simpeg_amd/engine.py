"""Device-side engine of the nodal DC simulation: one libsimpeg_dc handle per mesh.

Batches of node vectors are torch tensors of shape (ncol, ld) (column-major in the C ABI's
sense: column c starts at c*ld), float64 unless stated.  Every call runs on torch's current
stream of the engine's device.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .mesh import robin_face_weights


@dataclass
class SolveResult:
    outer: int
    iters: int
    max_rel_res: float
    converged: bool
    fp64_iters: int = 0


@dataclass
class GatherCSR:
    """d[i] = sum_e w[e] u_flat[idx[e]] (device arrays)."""
    nrow: int
    rowptr: torch.Tensor
    idx: torch.Tensor
    w: torch.Tensor


@dataclass
class ScatterCSR:
    """out_flat[dst[i]] = sum_e w[e] v[src[e]] (device arrays)."""
    nrow: int
    dst: torch.Tensor
    rowptr: torch.Tensor
    src: torch.Tensor
    w: torch.Tensor


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class NodalEngine:
    """Owns the C handle (mesh, multigrid hierarchy, solver workspace)."""

    def __init__(self, widths, origin, bc_type="Robin", surface_faces=None, device=None,
                 coarse_max=0, params=None):
        self.lib = _lib.load()
        if not torch.cuda.is_available():
            raise RuntimeError("no ROCm GPU is visible to torch (the nodal DC solver runs on the GPU)")
        self.device = torch.device("cuda", torch.cuda.current_device() if device is None else
                                   torch.device(device).index or 0)
        self.widths = [np.ascontiguousarray(h, dtype=np.float64) for h in widths]
        self.origin = np.asarray(origin, dtype=float)
        self.shape_cells = tuple(len(h) for h in self.widths)
        nx, ny, nz = self.shape_cells
        self.bc = _lib.SDC_BC_NEUMANN if bc_type == "Neumann" else _lib.SDC_BC_ROBIN
        d = _lib.MeshDesc()
        d.nx, d.ny, d.nz = nx, ny, nz
        d.hx, d.hy, d.hz = _dp(self.widths[0]), _dp(self.widths[1]), _dp(self.widths[2])
        d.bc_type = self.bc
        keep = []
        if self.bc == _lib.SDC_BC_ROBIN:
            beta = robin_face_weights(self.widths, self.origin, surface_faces)
            for key in ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax"):
                arr = np.ascontiguousarray(beta[key], dtype=np.float64)
                keep.append(arr)
                setattr(d, f"beta_{key}", _dp(arr))
            self.alpha = beta["alpha"]
        else:
            self.alpha = None
        d.coarse_max = int(coarse_max)
        d.device = self.device.index
        h = C.c_void_p()
        with torch.cuda.device(self.device):
            _lib.check(self.lib.sdc_create(C.byref(d), C.byref(h)), "sdc_create")
        self.h = h
        self.ld = int(self.lib.sdc_ld(h))
        self.nN = int(self.lib.sdc_num_nodes(h))
        self.nC = int(self.lib.sdc_num_cells(h))
        self.levels = int(self.lib.sdc_num_levels(h))
        self.params = _lib.Params()
        _lib.check(self.lib.sdc_default_params(C.byref(self.params)))
        if params:
            self.set_params(**params)
        self._sigma = None

    # ------------------------------------------------------------------ lifecycle
    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.sdc_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, **kw):
        for k, v in kw.items():
            if not hasattr(self.params, k):
                raise ValueError(f"unknown solver parameter {k!r}")
            setattr(self.params, k, v)
        _lib.check(self.lib.sdc_set_params(self.h, C.byref(self.params)), "sdc_set_params")

    def _stream(self):
        return _lib.stream_ptr(self.device)

    # ------------------------------------------------------------------ buffers
    def block(self, ncol, dtype=torch.float64, zero=True):
        f = torch.zeros if zero else torch.empty
        return f((max(int(ncol), 1), self.ld), dtype=dtype, device=self.device)

    def to_block(self, arr: np.ndarray, dtype=torch.float64):
        """(nN, ncol) host array -> (ncol, ld) device block."""
        arr = np.asarray(arr)
        if arr.ndim == 1:
            arr = arr[:, None]
        b = self.block(arr.shape[1], dtype)
        b[:, : self.nN] = torch.from_numpy(np.ascontiguousarray(arr.T)).to(self.device, dtype)
        return b

    def from_block(self, b, ncol=None) -> np.ndarray:
        """(ncol, ld) device block -> (nN, ncol) host array."""
        ncol = b.shape[0] if ncol is None else ncol
        return b[:ncol, : self.nN].T.contiguous().cpu().numpy()

    # ------------------------------------------------------------------ operator
    def set_sigma(self, sigma):
        if isinstance(sigma, torch.Tensor):
            s = sigma.to(self.device, torch.float64)
        else:
            s = torch.as_tensor(np.asarray(sigma, dtype=np.float64)).to(self.device)
        if s.numel() != self.nC:
            raise ValueError(f"sigma must have {self.nC} entries, got {s.numel()}")
        s = s.contiguous()
        _lib.check(self.lib.sdc_set_sigma(self.h, _lib.ptr(s), self._stream()), "sdc_set_sigma")
        self._sigma = s

    def set_sigma_sets(self, sigmas):
        """Build one operator per row of ``sigmas`` (nset, nC) on this mesh (operator sets, for
        solve_sets)."""
        if isinstance(sigmas, torch.Tensor):
            s = sigmas.to(self.device, torch.float64)
        else:
            s = torch.as_tensor(np.asarray(sigmas, dtype=np.float64)).to(self.device)
        if s.ndim != 2 or s.shape[1] != self.nC:
            raise ValueError(f"sigmas must have shape (nset, {self.nC}), got {tuple(s.shape)}")
        s = s.contiguous()
        _lib.check(self.lib.sdc_set_sigma_sets(self.h, int(s.shape[0]), _lib.ptr(s),
                                               self._stream()), "sdc_set_sigma_sets")
        self._sigma = s
        return s

    def solve_sets(self, q, colset, ncol=None, mode="mixed", tol=1e-10, inner_tol=1e-5,
                   max_outer=8, max_iter=500, u=None, raise_on_fail=True):
        """solve() with column c against operator set colset[c] (sets from set_sigma_sets; every
        aligned group of 8 columns must use one set)."""
        ncol = q.shape[0] if ncol is None else ncol
        cs = np.ascontiguousarray(np.asarray(colset, dtype=np.int32).ravel())
        if cs.size < ncol:
            raise ValueError(f"colset has {cs.size} entries, {ncol} needed")
        if u is None:
            u = self.block(q.shape[0], zero=False)
        o = _lib.SolveOpts()
        o.mode = {"fp64": _lib.SDC_MODE_FP64, "mixed": _lib.SDC_MODE_MIXED,
                  "mixed_ir": _lib.SDC_MODE_MIXED_IR}[mode]
        o.tol, o.inner_tol, o.max_outer, o.max_iter = tol, inner_tol, max_outer, max_iter
        st = _lib.SolveStats()
        rc = self.lib.sdc_solve_sets(self.h, _lib.ptr(q), _lib.ptr(u), ncol,
                                     cs.ctypes.data_as(C.c_void_p), C.byref(o), C.byref(st),
                                     self._stream())
        res = SolveResult(st.outer, st.iters, st.max_rel_res, bool(st.converged), st.fp64_iters)
        if rc != _lib.SDC_ERR_NOCONV or raise_on_fail:
            _lib.check(rc, "sdc_solve_sets")
        return u, res

    def set_coefficients(self, cx, cy, cz, bd):
        """Set the finest-level operator from its edge conductances (+x/+y/+z edge of every node)
        and the remaining diagonal instead of from sigma (host or device arrays of nN)."""
        arrs = []
        for a in (cx, cy, cz, bd):
            t = a if isinstance(a, torch.Tensor) else torch.as_tensor(np.asarray(a, np.float64))
            t = t.to(self.device, torch.float64).contiguous()
            if t.numel() != self.nN:
                raise ValueError(f"coefficient arrays must have {self.nN} entries, got {t.numel()}")
            arrs.append(t)
        _lib.check(self.lib.sdc_set_coefficients(self.h, *[_lib.ptr(t) for t in arrs],
                                                 self._stream()), "sdc_set_coefficients")
        self._coef = arrs

    def apply(self, x, ncol=None, dtype=None):
        ncol = x.shape[0] if ncol is None else ncol
        y = torch.empty_like(x)
        dt = _lib.SDC_F64 if x.dtype == torch.float64 else _lib.SDC_F32
        _lib.check(self.lib.sdc_apply(self.h, dt, _lib.ptr(x), _lib.ptr(y), ncol, self._stream()),
                   "sdc_apply")
        return y

    def vcycle(self, b, ncol=None):
        ncol = b.shape[0] if ncol is None else ncol
        z = torch.zeros_like(b)
        dt = _lib.SDC_F64 if b.dtype == torch.float64 else _lib.SDC_F32
        _lib.check(self.lib.sdc_vcycle(self.h, dt, _lib.ptr(b), _lib.ptr(z), ncol, self._stream()),
                   "sdc_vcycle")
        return z

    def solve(self, q, ncol=None, mode="mixed", tol=1e-10, inner_tol=1e-5, max_outer=8,
              max_iter=500, u=None, raise_on_fail=True):
        ncol = q.shape[0] if ncol is None else ncol
        if u is None:
            u = self.block(q.shape[0], zero=False)
        o = _lib.SolveOpts()
        o.mode = {"fp64": _lib.SDC_MODE_FP64, "mixed": _lib.SDC_MODE_MIXED,
                  "mixed_ir": _lib.SDC_MODE_MIXED_IR}[mode]
        o.tol, o.inner_tol, o.max_outer, o.max_iter = tol, inner_tol, max_outer, max_iter
        st = _lib.SolveStats()
        rc = self.lib.sdc_solve(self.h, _lib.ptr(q), _lib.ptr(u), ncol, C.byref(o), C.byref(st),
                                self._stream())
        res = SolveResult(st.outer, st.iters, st.max_rel_res, bool(st.converged), st.fp64_iters)
        if rc != _lib.SDC_ERR_NOCONV or raise_on_fail:
            _lib.check(rc, "sdc_solve")
        return u, res

    def profile(self, tags):
        """Bracket launches of kernel classes with events: a name ("smooth", "apply", ...), a
        list of names, "all", or None (off)."""
        if tags is None:
            names = []
        elif tags == "all":
            names = list(_lib.PROF_TAGS)
        elif isinstance(tags, str):
            names = [tags]
        else:
            names = list(tags)
        mask = 0
        for n in names:
            mask |= 1 << _lib.PROF_TAGS[n]
        _lib.check(self.lib.sdc_profile(self.h, mask), "sdc_profile")

    def profile_read(self, tag):
        """(launches, total ms, total algorithmic bytes) of one kernel class since profile() or
        the last read of that class."""
        n, ms, by = C.c_int64(), C.c_double(), C.c_double()
        _lib.check(self.lib.sdc_profile_read(self.h, _lib.PROF_TAGS[tag], C.byref(n), C.byref(ms),
                                             C.byref(by)), "sdc_profile_read")
        return int(n.value), float(ms.value), float(by.value)

    def jvec_rhs(self, dsig, u, ncol=None, out=None):
        ncol = u.shape[0] if ncol is None else ncol
        if isinstance(dsig, torch.Tensor):
            ds = dsig.to(self.device, torch.float64).contiguous()
        else:
            ds = torch.as_tensor(np.asarray(dsig, dtype=np.float64)).to(self.device).contiguous()
        if out is None:
            out = torch.empty_like(u)
        elif out.shape[0] < ncol or out.shape[1] != self.ld or not out.is_contiguous():
            raise ValueError("out must be a contiguous (>= ncol, ld) block")
        _lib.check(self.lib.sdc_jvec_rhs(self.h, _lib.ptr(ds), _lib.ptr(u), _lib.ptr(out), ncol,
                                         self._stream()), "sdc_jvec_rhs")
        return out

    def _check_pairing(self, u, lam, ncol, ucol):
        """Column j of lam pairs with u column ucol[j] (j when ucol is None): every paired column
        must exist, since the kernels index u and lam without bounds."""
        if ncol < 1 or ncol > lam.shape[0]:
            raise ValueError(f"ncol={ncol} outside the {lam.shape[0]} columns of lam")
        if ucol is None:
            if ncol > u.shape[0]:
                raise ValueError(f"lam has {ncol} columns but u only {u.shape[0]}")
            return None
        uc = np.asarray(ucol.cpu() if isinstance(ucol, torch.Tensor) else ucol).ravel()
        if uc.size < ncol:
            raise ValueError(f"ucol has {uc.size} entries, {ncol} needed")
        if uc.size and (uc[:ncol].min() < 0 or uc[:ncol].max() >= u.shape[0]):
            raise ValueError(f"ucol entries must lie in [0, {u.shape[0]})")
        return torch.as_tensor(uc[:ncol].astype(np.int32)).to(self.device).contiguous()

    def contract(self, u, lam, ncol=None, ucol=None, per_column=False, out=None):
        """sum over columns (or, per_column, one row per column) of the adjoint contraction;
        ``out``: a contiguous float64 device tensor of that shape to write into."""
        ncol = lam.shape[0] if ncol is None else ncol
        ucol = self._check_pairing(u, lam, ncol, ucol)
        shape = (ncol, self.nC) if per_column else (self.nC,)
        if out is None:
            out = torch.empty(shape, dtype=torch.float64, device=self.device)
        elif (tuple(out.shape) != shape or out.dtype != torch.float64 or not out.is_contiguous()
              or out.device != torch.device(self.device)):
            raise ValueError(f"out must be a contiguous float64 {shape} tensor on {self.device}")
        uc = ucol
        _lib.check(self.lib.sdc_jtvec_contract(self.h, _lib.ptr(u), _lib.ptr(lam), ncol,
                                               _lib.ptr(uc) if uc is not None else None,
                                               1 if per_column else 0, _lib.ptr(out), self.nC,
                                               self._stream()), "sdc_jtvec_contract")
        return out

    def jtj_diag_accumulate(self, u, lam, ncol, ucol, wsq, out):
        """out[c] += sum_j wsq[j] g_j[c]^2 over ncol adjoint columns (device arrays)."""
        uc = self._check_pairing(u, lam, ncol, ucol)
        if uc is None:
            uc = torch.arange(ncol, dtype=torch.int32, device=self.device)
        w = torch.as_tensor(np.asarray(wsq, dtype=np.float64)).to(self.device).contiguous()
        _lib.check(self.lib.sdc_jtj_diag_accumulate(self.h, _lib.ptr(u), _lib.ptr(lam), ncol,
                                                    _lib.ptr(uc), _lib.ptr(w), _lib.ptr(out),
                                                    self._stream()), "sdc_jtj_diag_accumulate")
        return out

    # ------------------------------------------------------------------ projections
    def gather_csr(self, rowptr, idx, w) -> GatherCSR:
        dev = self.device
        return GatherCSR(len(rowptr) - 1,
                         torch.as_tensor(np.asarray(rowptr, dtype=np.int32)).to(dev),
                         torch.as_tensor(np.asarray(idx, dtype=np.int64)).to(dev),
                         torch.as_tensor(np.asarray(w, dtype=np.float64)).to(dev))

    def scatter_csr(self, dst, rowptr, src, w) -> ScatterCSR:
        dev = self.device
        return ScatterCSR(len(dst),
                          torch.as_tensor(np.asarray(dst, dtype=np.int64)).to(dev),
                          torch.as_tensor(np.asarray(rowptr, dtype=np.int32)).to(dev),
                          torch.as_tensor(np.asarray(src, dtype=np.int32)).to(dev),
                          torch.as_tensor(np.asarray(w, dtype=np.float64)).to(dev))

    def gather(self, g: GatherCSR, u):
        d = torch.empty((max(g.nrow, 1),), dtype=torch.float64, device=self.device)
        _lib.check(self.lib.sdc_gather(self.h, g.nrow, _lib.ptr(g.rowptr), _lib.ptr(g.idx),
                                       _lib.ptr(g.w), _lib.ptr(u), _lib.ptr(d), self._stream()),
                   "sdc_gather")
        return d[: g.nrow]

    def scatter(self, s: ScatterCSR, v, ncol, out=None):
        """P^T v into ncol columns (zeroed first) of a new block, or of ``out``."""
        v = torch.as_tensor(v, dtype=torch.float64).to(self.device).contiguous()
        if out is None:
            out = self.block(ncol, zero=False)
        elif out.shape[0] < ncol or out.shape[1] != self.ld or not out.is_contiguous():
            raise ValueError("out must be a contiguous (>= ncol, ld) block")
        _lib.check(self.lib.sdc_scatter(self.h, s.nrow, _lib.ptr(s.dst), _lib.ptr(s.rowptr),
                                        _lib.ptr(s.src), _lib.ptr(s.w), _lib.ptr(v), _lib.ptr(out),
                                        ncol, self._stream()), "sdc_scatter")
        return out


def download(x):
    """A device tensor as a host array: through page-locked memory (a DMA copy on the current
    stream; torch's host allocator keeps the block for the next call) when x is on a GPU."""
    if not x.is_cuda:
        return x.numpy()
    host = torch.empty(x.shape, dtype=x.dtype, pin_memory=True)
    host.copy_(x, non_blocking=True)
    torch.cuda.current_stream(x.device).synchronize()
    return host.numpy()


def coo_to_gather(rows, flat, w, nrow):
    """COO (row, flat index, weight) -> CSR arrays (duplicates kept)."""
    order = np.lexsort((flat, rows))
    rows, flat, w = rows[order], flat[order], w[order]
    rowptr = np.zeros(nrow + 1, dtype=np.int64)
    np.add.at(rowptr, rows + 1, 1)
    return np.cumsum(rowptr), flat, w


def coo_to_scatter(dest, src, w):
    """COO (destination flat index, source index, weight) -> grouped by destination."""
    order = np.lexsort((src, dest))
    dest, src, w = dest[order], src[order], w[order]
    uniq, start = np.unique(dest, return_index=True)
    rowptr = np.r_[start, dest.size]
    return uniq, rowptr, src, w
