"""FP64 CPU restatement of SimPEG's 3D nodal DC simulation (oracle; test infrastructure only).

Reference files followed (paths relative to the simpeg repository root):

* ``simpeg/electromagnetics/static/resistivity/simulation.py``
  - ``BaseDCSimulation.fields`` 100-113, ``getJ`` 115-121, ``dpred`` 123-135,
    ``getJtJdiag`` 137-154, ``Jvec`` 156-187, ``Jtvec``/``_Jtvec`` 189-263,
    ``getSourceTerm`` 265-289
  - ``Simulation3DNodal.getA`` 523-555, ``getADeriv`` 557-571, ``setBC`` 573-644
* ``simpeg/base/pde_simulation.py`` ``_inner_mat_mul_op`` 15-83, ``MeSigma`` 150-162,
  ``MeSigmaDeriv`` 296-323
* ``simpeg/electromagnetics/static/resistivity/sources.py`` ``BaseSrc.eval`` 64-88
* ``simpeg/electromagnetics/static/resistivity/receivers.py`` ``BaseRx.eval`` 156-192,
  ``evalDeriv`` 194-244, ``Dipole.getP`` 379-424, ``Pole.getP`` 461-498
* ``simpeg/simulation.py`` ``BaseSimulation.dpred`` 161-196
* ``simpeg/maps/_property_maps.py`` ``ExpMap`` 18-112

The discretize operators (not vendored in the reference) are restated from their
published definitions: Fortran (x-fastest) ordering of nodes/cells/edges/faces,
kron-product construction of the nodal gradient and averaging operators, the
isotropic edge inner product ``M_e = 3 diag(aveE2CC^T (V sigma))`` and cell-to-face
averaging with constant extrapolation on boundary faces.
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp
from scipy.sparse.linalg import splu


# --------------------------------------------------------------------------- mesh


def unpack_widths(value):
    """discretize ``unpack_widths``: scalars, (h, n) and (h, n, factor) tuples.

    (h, n, f) expands to h*|f|**k for k = 1..n, reversed when f < 0.
    """
    if isinstance(value, np.ndarray):
        return np.asarray(value, dtype=float).ravel()
    out = []
    for v in value:
        if np.isscalar(v):
            out.append(float(v))
        elif isinstance(v, tuple) and len(v) == 2:
            out += [float(v[0])] * int(v[1])
        elif isinstance(v, tuple) and len(v) == 3:
            start, num, factor = float(v[0]), int(v[1]), float(v[2])
            pad = (np.abs(factor) ** (np.arange(num) + 1)) * start
            if factor < 0:
                pad = pad[::-1]
            out += pad.tolist()
        else:
            raise ValueError("widths must be scalars, (h, n) or (h, n, factor) tuples")
    return np.asarray(out, dtype=float)


def _origin_axis(h, code):
    if isinstance(code, str):
        if code == "C":
            return -h.sum() / 2.0
        if code == "N":
            return -h.sum()
        if code == "0":
            return 0.0
        raise ValueError(f"unknown origin code {code!r}")
    return float(code)


def _speye(n):
    return sp.identity(n, format="csr")


def _kron3(a, b, c):
    return sp.kron(a, sp.kron(b, c, format="csr"), format="csr")


def _ddx(n):
    """(n, n+1) difference operator u[i+1]-u[i]."""
    return sp.diags([-np.ones(n), np.ones(n)], [0, 1], shape=(n, n + 1), format="csr")


def _av(n):
    """(n, n+1) node->cell average."""
    return sp.diags([0.5 * np.ones(n), 0.5 * np.ones(n)], [0, 1], shape=(n, n + 1), format="csr")


def _av_extrap(n):
    """(n+1, n) cell->node average with constant extrapolation at both ends."""
    a = sp.diags([0.5 * np.ones(n), 0.5 * np.ones(n)], [-1, 0], shape=(n + 1, n), format="lil")
    a[0, 0] = 1.0
    a[n, n - 1] = 1.0
    return a.tocsr()


def _boundary_bool(shape, bdir):
    is_b = np.zeros(shape, dtype=bool, order="F")
    if "x" in bdir:
        is_b[[0, -1], :, :] = True
    if "y" in bdir:
        is_b[:, [0, -1], :] = True
    if "z" in bdir:
        is_b[:, :, [0, -1]] = True
    return is_b.reshape(-1, order="F")


class OracleMesh:
    """Restatement of the parts of ``discretize.TensorMesh`` (3D) the path uses."""

    def __init__(self, h, origin=None):
        self.h = [unpack_widths(hh) for hh in h]
        if len(self.h) != 3:
            raise ValueError("the oracle handles 3D tensor meshes only")
        if origin is None:
            origin = ["0", "0", "0"]
        if isinstance(origin, str):
            origin = list(origin)
        self.origin = np.array([_origin_axis(hh, o) for hh, o in zip(self.h, origin)])
        self.shape_cells = tuple(len(hh) for hh in self.h)

    # geometry --------------------------------------------------------------
    @property
    def nodes_x(self):
        return self.origin[0] + np.r_[0.0, np.cumsum(self.h[0])]

    @property
    def nodes_y(self):
        return self.origin[1] + np.r_[0.0, np.cumsum(self.h[1])]

    @property
    def nodes_z(self):
        return self.origin[2] + np.r_[0.0, np.cumsum(self.h[2])]

    def _cc(self, a):
        n = [self.nodes_x, self.nodes_y, self.nodes_z][a]
        return 0.5 * (n[1:] + n[:-1])

    @property
    def shape_nodes(self):
        return tuple(n + 1 for n in self.shape_cells)

    @property
    def n_cells(self):
        return int(np.prod(self.shape_cells))

    @property
    def n_nodes(self):
        return int(np.prod(self.shape_nodes))

    @staticmethod
    def _ndgrid(x, y, z):
        X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
        return np.c_[X.reshape(-1, order="F"), Y.reshape(-1, order="F"), Z.reshape(-1, order="F")]

    @property
    def nodes(self):
        return self._ndgrid(self.nodes_x, self.nodes_y, self.nodes_z)

    @property
    def cell_centers(self):
        return self._ndgrid(self._cc(0), self._cc(1), self._cc(2))

    @property
    def cell_volumes(self):
        hx, hy, hz = self.h
        return np.outer(np.outer(hx, hy).reshape(-1, order="F"), hz).reshape(-1, order="F")

    # operators ---------------------------------------------------------------
    @property
    def nodal_gradient(self):
        nx, ny, nz = self.shape_cells
        hx, hy, hz = self.h
        gx = _kron3(_speye(nz + 1), _speye(ny + 1), sp.diags(1.0 / hx) @ _ddx(nx))
        gy = _kron3(_speye(nz + 1), sp.diags(1.0 / hy) @ _ddx(ny), _speye(nx + 1))
        gz = _kron3(sp.diags(1.0 / hz) @ _ddx(nz), _speye(ny + 1), _speye(nx + 1))
        return sp.vstack([gx, gy, gz], format="csr")

    @property
    def average_edge_to_cell(self):
        nx, ny, nz = self.shape_cells
        ex = _kron3(_av(nz), _av(ny), _speye(nx))
        ey = _kron3(_av(nz), _speye(ny), _av(nx))
        ez = _kron3(_speye(nz), _av(ny), _av(nx))
        return (1.0 / 3.0) * sp.hstack([ex, ey, ez], format="csr")

    @property
    def edge_cell_weights(self):
        """A_v = 3 aveE2CC^T diag(V): M_e(sigma) = diag(A_v sigma)."""
        return (3.0 * self.average_edge_to_cell.T @ sp.diags(self.cell_volumes)).tocsr()

    def get_edge_inner_product(self, sigma):
        return sp.diags(self.edge_cell_weights @ sigma, format="csr")

    @property
    def shape_faces(self):
        nx, ny, nz = self.shape_cells
        return [(nx + 1, ny, nz), (nx, ny + 1, nz), (nx, ny, nz + 1)]

    @property
    def average_cell_to_face(self):
        nx, ny, nz = self.shape_cells
        fx = _kron3(_speye(nz), _speye(ny), _av_extrap(nx))
        fy = _kron3(_speye(nz), _av_extrap(ny), _speye(nx))
        fz = _kron3(_av_extrap(nz), _speye(ny), _speye(nx))
        return sp.vstack([fx, fy, fz], format="csr")

    @property
    def average_node_to_face(self):
        nx, ny, nz = self.shape_cells
        fx = _kron3(_av(nz), _av(ny), _speye(nx + 1))
        fy = _kron3(_av(nz), _speye(ny + 1), _av(nx))
        fz = _kron3(_speye(nz + 1), _av(ny), _av(nx))
        return sp.vstack([fx, fy, fz], format="csr")

    @property
    def face_areas(self):
        hx, hy, hz = self.h
        nx, ny, nz = self.shape_cells
        ax = np.outer(np.ones(nx + 1), np.outer(hy, hz).reshape(-1, order="F")).reshape(-1, order="F")
        ay = np.outer(np.outer(hx, np.ones(ny + 1)).reshape(-1, order="F"), hz).reshape(-1, order="F")
        az = np.outer(np.outer(hx, hy).reshape(-1, order="F"), np.ones(nz + 1)).reshape(-1, order="F")
        return np.r_[ax, ay, az]

    @property
    def _is_boundary_face(self):
        sx, sy, sz = self.shape_faces
        return np.r_[_boundary_bool(sx, "x"), _boundary_bool(sy, "y"), _boundary_bool(sz, "z")]

    @property
    def project_face_to_boundary_face(self):
        is_b = self._is_boundary_face
        return sp.identity(is_b.size, format="csr")[is_b]

    @property
    def boundary_faces(self):
        cx, cy, cz = self._cc(0), self._cc(1), self._cc(2)
        fx = self._ndgrid(self.nodes_x[[0, -1]], cy, cz)
        fy = self._ndgrid(cx, self.nodes_y[[0, -1]], cz)
        fz = self._ndgrid(cx, cy, self.nodes_z[[0, -1]])
        return np.r_[fx, fy, fz]

    @property
    def boundary_face_outward_normals(self):
        nx, ny, nz = self.shape_cells
        sgn = np.array([-1.0, 1.0])
        nxv = np.zeros((2 * ny * nz, 3))
        nxv[:, 0] = np.tile(sgn, ny * nz)
        nyv = np.zeros((2 * nx * nz, 3))
        nyv[:, 1] = np.tile(np.repeat(sgn, nx), nz)
        nzv = np.zeros((2 * nx * ny, 3))
        nzv[:, 2] = np.repeat(sgn, nx * ny)
        return np.r_[nxv, nyv, nzv]

    # interpolation -----------------------------------------------------------
    @staticmethod
    def _interp1d(xn, xr):
        if np.any(xr < xn[0] - 1e-9 * abs(xn[-1] - xn[0])) or np.any(xr > xn[-1] + 1e-9 * abs(xn[-1] - xn[0])):
            raise ValueError("points outside of mesh")
        i = np.searchsorted(xn, xr, side="right") - 1
        i = np.clip(i, 0, xn.size - 2)
        w1 = (xr - xn[i]) / (xn[i + 1] - xn[i])
        return i, 1.0 - w1, w1

    def get_interpolation_matrix(self, locs):
        """Trilinear node interpolation matrix (n_loc, n_nodes)."""
        locs = np.atleast_2d(np.asarray(locs, dtype=float))
        n1x, n1y, _ = self.shape_nodes
        ix, wx0, wx1 = self._interp1d(self.nodes_x, locs[:, 0])
        iy, wy0, wy1 = self._interp1d(self.nodes_y, locs[:, 1])
        iz, wz0, wz1 = self._interp1d(self.nodes_z, locs[:, 2])
        rows, cols, vals = [], [], []
        for a, wa in ((0, wx0), (1, wx1)):
            for b, wb in ((0, wy0), (1, wy1)):
                for c, wc in ((0, wz0), (1, wz1)):
                    rows.append(np.arange(locs.shape[0]))
                    cols.append((ix + a) + n1x * ((iy + b) + n1y * (iz + c)))
                    vals.append(wa * wb * wc)
        P = sp.csr_matrix(
            (np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
            shape=(locs.shape[0], self.n_nodes),
        )
        P.eliminate_zeros()
        return P


# --------------------------------------------------------------------------- maps


class OracleIdentityMap:
    """``maps.IdentityMap`` (``maps/_base.py:26-256``)."""

    def transform(self, m):
        return np.asarray(m, dtype=float).copy()

    def deriv(self, m):
        return sp.identity(np.size(m), format="csr")


class OracleExpMap:
    """``maps.ExpMap`` (``maps/_property_maps.py:47-48, 105-108``)."""

    def transform(self, m):
        return np.exp(np.asarray(m, dtype=float))

    def deriv(self, m):
        return sp.diags(np.exp(np.asarray(m, dtype=float)), format="csr")


# --------------------------------------------------------------------------- simulation


def robin_boundary_matrix(mesh: OracleMesh, surface_faces=None):
    """``Simulation3DNodal.setBC`` Robin branch (``resistivity/simulation.py:579-644``).

    Returns ``_AvgBC`` (n_nodes x n_cells) and the boundary-face alpha.
    """
    bfaces = mesh.boundary_faces
    normals = mesh.boundary_face_outward_normals
    n_bf = len(bfaces)
    alpha = np.zeros(n_bf)
    nodes = mesh.nodes
    middle = np.median(nodes, axis=0)
    top_v = np.max(nodes[:, -1])
    source_point = np.r_[middle[:-1], top_v]
    r_vec = bfaces - source_point
    r = np.linalg.norm(r_vec, axis=-1)
    r_hat = r_vec / r[:, None]
    r_dot_n = np.einsum("ij,ij->i", r_hat, normals)
    if surface_faces is None:
        # tensor branch (:618-627): the z-face boundary flags in Fortran order, top plane last
        shape_fz = mesh.shape_faces[2]
        is_b = _boundary_bool(shape_fz, "xyz")
        is_t = np.zeros(shape_fz, dtype=bool, order="F")
        is_t[:, :, -1] = True
        is_t = is_t.reshape(-1, order="F")[is_b]
        not_top = np.ones(n_bf, dtype=bool)
        not_top[-len(is_t):] = ~is_t
    else:
        not_top = ~np.asarray(surface_faces, dtype=bool)
    alpha[not_top] = (r_dot_n / r)[not_top]
    P_bf = mesh.project_face_to_boundary_face
    avg_n2fb = P_bf @ mesh.average_node_to_face
    avg_cc2fb = P_bf @ mesh.average_cell_to_face
    avg_cc2fb = sp.diags(alpha * (P_bf @ mesh.face_areas)) @ avg_cc2fb
    return (avg_n2fb.T @ avg_cc2fb).tocsr(), alpha


class OracleDC:
    """CPU FP64 Simulation3DNodal restatement.

    ``sources`` is a list of ``(a_locs (k,3), currents (k,), receivers)`` where
    ``receivers`` is a list of dicts ``{"m": (n,3), "n": (n,3) or None,
    "factor": None or (n,)}`` (``factor`` = 1/geometric factor for apparent
    resistivity data).  Data are ordered source-major, then receiver, then location
    (``simpeg/survey.py:592-627``).
    """

    def __init__(self, mesh: OracleMesh, sources, mapping=None, bc_type="Robin", surface_faces=None,
                 param="sigma"):
        if bc_type not in ("Robin", "Mixed", "Neumann"):
            raise ValueError("bc_type must be 'Robin' or 'Neumann' for the nodal formulation")
        if param not in ("sigma", "rho"):
            raise ValueError("param must be 'sigma' (sigmaMap) or 'rho' (rhoMap)")
        self.mesh = mesh
        self.sources = sources
        self.mapping = mapping if mapping is not None else OracleExpMap()
        # "rho": the mapping gives the resistivity (rhoMap); sigma = 1/rho through the reciprocal
        # relation of props.py:47-56, 139-157
        self.param = param
        self.bc_type = "Neumann" if bc_type == "Neumann" else "Robin"
        self.G = mesh.nodal_gradient
        self.Av = mesh.edge_cell_weights
        if self.bc_type == "Robin":
            self.AvgBC, self.alpha = robin_boundary_matrix(mesh, surface_faces)
        else:
            self.AvgBC, self.alpha = None, None
        self._P = []
        for _, _, rxs in sources:
            plist = []
            for rx in rxs:
                P = mesh.get_interpolation_matrix(rx["m"])
                if rx.get("n") is not None:
                    P = P - mesh.get_interpolation_matrix(rx["n"])
                if rx.get("factor") is not None:
                    P = sp.diags(rx["factor"]) @ P
                plist.append(P.tocsr())
            self._P.append(plist)
        self.model = None
        self._lu = None

    # ---- model / operator
    def set_model(self, m):
        m = np.asarray(m, dtype=float)
        if self.model is None or not np.allclose(self.model, m):
            self.model = m.copy()
            if self.param == "rho":
                self.sigma = 1.0 / self.mapping.transform(m)
                self.sigma_deriv = (sp.diags(-self.sigma * self.sigma) @ self.mapping.deriv(m)).tocsr()
            else:
                self.sigma = self.mapping.transform(m)
                self.sigma_deriv = self.mapping.deriv(m)
            self._lu = None

    def getA(self, sigma=None):
        """``Simulation3DNodal.getA`` (523-555)."""
        s = self.sigma if sigma is None else sigma
        A = (self.G.T @ sp.diags(self.Av @ s) @ self.G).tocsr()
        if self.bc_type == "Neumann":
            A = A.tolil()
            A[0, :] = 0.0
            A[0, 0] = 1.0
            A = A.tocsr()
        else:
            A = A + sp.diags(self.AvgBC @ s, format="csr")
        return A

    def getADeriv(self, u, v, adjoint=False):
        """``Simulation3DNodal.getADeriv`` (557-571) with ``MeSigmaDeriv`` (pde_simulation.py:296-321).

        The Robin term is added only for a conductivity model (``sigmaMap is not None``, :566): a
        resistivity model (``param="rho"``, rhoMap) leaves it out, as the reference does."""
        Gu = self.G @ u
        M = self.Av @ self.sigma_deriv
        if not adjoint:
            out = self.G.T @ (Gu * (M @ v))
        else:
            Gv = self.G @ v
            if Gv.ndim > 1:
                out = M.T @ (Gu[:, None] * Gv)
            else:
                out = M.T @ (Gu * Gv)
        if self.bc_type != "Neumann" and self.param == "sigma":
            MBC = self.AvgBC @ self.sigma_deriv
            if not adjoint:
                out = out + u * (MBC @ v)
            else:
                out = out + (MBC.T @ (u[:, None] * v) if np.ndim(v) > 1 else MBC.T @ (u * v))
        return out

    def source_term(self):
        """``getSourceTerm`` (265-289) with ``BaseSrc.eval`` (sources.py:64-88)."""
        q = np.zeros((self.mesh.n_nodes, len(self.sources)), order="F")
        for i, (locs, cur, _) in enumerate(self.sources):
            P = self.mesh.get_interpolation_matrix(locs).toarray()
            q[:, i] = np.sum(np.asarray(cur, dtype=float)[:, None] * P, axis=0)
        return q

    def factor(self):
        if self._lu is None:
            self._lu = splu(self.getA().tocsc())
        return self._lu

    def solve(self, rhs):
        return self.factor().solve(np.asarray(rhs, dtype=float))

    # ---- public API
    def fields(self, m):
        """``BaseDCSimulation.fields`` (100-113): (n_nodes, n_src) potentials."""
        self.set_model(m)
        self.factor()
        return self.solve(self.source_term())

    @property
    def nD(self):
        return sum(P.shape[0] for plist in self._P for P in plist)

    def dpred(self, m, u=None):
        """``BaseSimulation.dpred`` (simulation.py:161-196)."""
        if u is None:
            u = self.fields(m)
        out = []
        for s, plist in enumerate(self._P):
            for P in plist:
                out.append(P @ u[:, s])
        return np.concatenate(out)

    def Jvec(self, m, v, u=None):
        """``BaseDCSimulation.Jvec`` (156-187)."""
        if u is None:
            u = self.fields(m)
        self.set_model(m)
        out = []
        for s, plist in enumerate(self._P):
            dA = self.getADeriv(u[:, s], v)
            du = self.solve(-dA)
            for P in plist:
                out.append(P @ du)
        return np.concatenate(out)

    def Jtvec(self, m, v, u=None):
        """``BaseDCSimulation._Jtvec`` (205-263), one adjoint solve per (src, rx) pair."""
        if u is None:
            u = self.fields(m)
        self.set_model(m)
        Jtv = np.zeros(np.size(m))
        i0 = 0
        for s, plist in enumerate(self._P):
            for P in plist:
                PTv = P.T @ v[i0:i0 + P.shape[0]]
                i0 += P.shape[0]
                lam = self.solve(PTv)
                Jtv += -self.getADeriv(u[:, s], lam, adjoint=True)
        return Jtv

    def getJ(self, m, u=None):
        """``getJ`` (115-121) via ``_Jtvec(v=None)`` (222-258): dense (nD, nP)."""
        if u is None:
            u = self.fields(m)
        self.set_model(m)
        cols = []
        for s, plist in enumerate(self._P):
            for P in plist:
                PT = P.toarray().T
                lam = self.solve(PT)
                cols.append(-self.getADeriv(u[:, s], lam, adjoint=True))
        return np.hstack(cols).T

    def getJtJdiag(self, m, W=None, u=None):
        """``getJtJdiag`` (137-154): sum_i W_ii^2 J_i^2."""
        J = self.getJ(m, u=u)
        w = np.ones(J.shape[0]) if W is None else np.asarray(W.diagonal()) ** 2
        return (w[:, None] * J * J).sum(axis=0)
