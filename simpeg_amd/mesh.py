"""3D tensor mesh with the discretize ``TensorMesh`` attributes the nodal DC path uses.

The simulation also accepts a real ``discretize.TensorMesh`` (or anything exposing
``h`` and ``origin``): only the cell widths and the origin are read from it.

Ordering follows discretize: nodes and cells are Fortran ordered (x fastest); the
boundary faces are the x faces (2, ny, nz), then the y faces (nx, 2, nz), then the z
faces (nx, ny, 2), each Fortran ordered.
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp


def unpack_widths(value):
    """Expand scalars, (h, n) and (h, n, factor) tuples into cell widths.

    (h, n, f) gives h*|f|**k for k = 1..n, reversed when f < 0 (discretize semantics).
    """
    if isinstance(value, np.ndarray):
        return np.asarray(value, dtype=float).ravel()
    out = []
    for v in value:
        if np.isscalar(v):
            out.append(float(v))
        elif isinstance(v, tuple) and len(v) == 2:
            out.extend([float(v[0])] * int(v[1]))
        elif isinstance(v, tuple) and len(v) == 3:
            h0, num, f = float(v[0]), int(v[1]), float(v[2])
            pad = h0 * np.abs(f) ** (np.arange(num) + 1.0)
            out.extend((pad[::-1] if f < 0 else pad).tolist())
        else:
            raise ValueError("cell widths must be scalars, (h, n) or (h, n, factor) tuples")
    return np.asarray(out, dtype=float)


def _origin(h, code):
    if isinstance(code, str):
        return {"C": -h.sum() / 2.0, "N": -h.sum(), "0": 0.0}[code]
    return float(code)


class TensorMesh:
    """Minimal 3D ``discretize.TensorMesh`` (widths, origin, ordering, a few operators)."""

    _meshType = "TENSOR"
    dim = 3

    def __init__(self, h, origin=None, x0=None):
        if origin is None:
            origin = x0
        self.h = [unpack_widths(hh) for hh in h]
        if len(self.h) != 3:
            raise NotImplementedError("only 3D tensor meshes are supported")
        if any((hh <= 0).any() or hh.size == 0 for hh in self.h):
            raise ValueError("cell widths must be positive")
        if origin is None:
            origin = "000"
        if isinstance(origin, str):
            origin = list(origin)
        self.origin = np.array([_origin(hh, o) for hh, o in zip(self.h, origin)], dtype=float)

    # sizes -------------------------------------------------------------------------------
    @property
    def shape_cells(self):
        return tuple(len(hh) for hh in self.h)

    @property
    def shape_nodes(self):
        return tuple(n + 1 for n in self.shape_cells)

    @property
    def n_cells(self):
        return int(np.prod(self.shape_cells))

    nC = n_cells

    @property
    def n_nodes(self):
        return int(np.prod(self.shape_nodes))

    nN = n_nodes

    @property
    def n_edges(self):
        nx, ny, nz = self.shape_cells
        return nx * (ny + 1) * (nz + 1) + (nx + 1) * ny * (nz + 1) + (nx + 1) * (ny + 1) * nz

    def __len__(self):
        return self.n_cells

    # coordinates -------------------------------------------------------------------------
    def _nodes_axis(self, a):
        return self.origin[a] + np.r_[0.0, np.cumsum(self.h[a])]

    @property
    def nodes_x(self):
        return self._nodes_axis(0)

    @property
    def nodes_y(self):
        return self._nodes_axis(1)

    @property
    def nodes_z(self):
        return self._nodes_axis(2)

    @property
    def cell_centers_x(self):
        n = self.nodes_x
        return 0.5 * (n[1:] + n[:-1])

    @property
    def cell_centers_y(self):
        n = self.nodes_y
        return 0.5 * (n[1:] + n[:-1])

    @property
    def cell_centers_z(self):
        n = self.nodes_z
        return 0.5 * (n[1:] + n[:-1])

    @staticmethod
    def _grid(x, y, z):
        X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
        return np.c_[X.ravel(order="F"), Y.ravel(order="F"), Z.ravel(order="F")]

    @property
    def nodes(self):
        return self._grid(self.nodes_x, self.nodes_y, self.nodes_z)

    @property
    def cell_centers(self):
        return self._grid(self.cell_centers_x, self.cell_centers_y, self.cell_centers_z)

    @property
    def cell_volumes(self):
        hx, hy, hz = self.h
        return (hx[:, None, None] * hy[None, :, None] * hz[None, None, :]).ravel(order="F")

    # operators (host, scipy) used by the derived field aliases --------------------------------
    @property
    def nodal_gradient(self):
        if getattr(self, "_grad", None) is None:
            nx, ny, nz = self.shape_cells

            def d(n, h):
                return sp.diags(1.0 / h) @ sp.diags([-np.ones(n), np.ones(n)], [0, 1], shape=(n, n + 1))

            I = lambda n: sp.identity(n, format="csr")  # noqa: E731
            k3 = lambda a, b, c: sp.kron(a, sp.kron(b, c))  # noqa: E731
            hx, hy, hz = self.h
            self._grad = sp.vstack(
                [k3(I(nz + 1), I(ny + 1), d(nx, hx)), k3(I(nz + 1), d(ny, hy), I(nx + 1)),
                 k3(d(nz, hz), I(ny + 1), I(nx + 1))],
                format="csr",
            )
        return self._grad

    @property
    def average_node_to_cell(self):
        nx, ny, nz = self.shape_cells

        def av(n):
            return sp.diags([0.5 * np.ones(n), 0.5 * np.ones(n)], [0, 1], shape=(n, n + 1))

        return sp.kron(av(nz), sp.kron(av(ny), av(nx))).tocsr()

    aveN2CC = average_node_to_cell


# ------------------------------------------------------------------------------------------------
# helpers shared by the simulation (geometry of any mesh exposing h / origin)


def mesh_widths(mesh):
    """Cell widths (hx, hy, hz) and origin of a tensor mesh (ours or discretize's)."""
    h = getattr(mesh, "h", None)
    if h is None or len(h) != 3:
        raise NotImplementedError("Simulation3DNodal on the GPU supports 3D tensor meshes only")
    origin = getattr(mesh, "origin", None)
    if origin is None:
        origin = getattr(mesh, "x0")
    return [np.asarray(hh, dtype=float) for hh in h], np.asarray(origin, dtype=float)


def node_interpolation(widths, origin, locs):
    """Trilinear interpolation of node values at points.

    Returns (idx, w): (n_loc, 8) node indices and weights (Fortran node numbering).  A point
    on a grid line takes weight 1 on that node.  Points outside the mesh raise ValueError.
    """
    locs = np.atleast_2d(np.asarray(locs, dtype=float))
    if locs.shape[1] != 3:
        raise ValueError("locations must have 3 columns (x, y, z)")
    n1 = [len(hh) + 1 for hh in widths]
    ii, w0, w1 = [], [], []
    for a in range(3):
        xn = origin[a] + np.r_[0.0, np.cumsum(widths[a])]
        xr = locs[:, a]
        span = xn[-1] - xn[0]
        if np.any(xr < xn[0] - 1e-9 * span) or np.any(xr > xn[-1] + 1e-9 * span):
            raise ValueError("points outside of mesh")
        i = np.clip(np.searchsorted(xn, xr, side="right") - 1, 0, xn.size - 2)
        t = (xr - xn[i]) / (xn[i + 1] - xn[i])
        ii.append(i)
        w0.append(1.0 - t)
        w1.append(t)
    idx = np.empty((locs.shape[0], 8), dtype=np.int64)
    w = np.empty((locs.shape[0], 8))
    m = 0
    for c in range(2):
        for b in range(2):
            for a in range(2):
                idx[:, m] = (ii[0] + a) + n1[0] * ((ii[1] + b) + n1[1] * (ii[2] + c))
                w[:, m] = (w1[0] if a else w0[0]) * (w1[1] if b else w0[1]) * (w1[2] if c else w0[2])
                m += 1
    return idx, w


def robin_face_weights(widths, origin, surface_faces=None):
    """Robin weights beta_f = alpha_f |f| / 4 on the six boundary planes.

    alpha_f = (r_hat . n_f) / r with r measured from (median node x, median node y, top z),
    and alpha_f = 0 on surface faces (by default the z = max plane), as in
    ``Simulation3DNodal.setBC`` (resistivity/simulation.py:586-644).

    Returns dict with keys xmin, xmax (ny*nz, index j + ny*k), ymin, ymax (nx*nz, i + nx*k),
    zmin, zmax (nx*ny, i + nx*j), plus ``alpha`` in discretize boundary-face order.
    """
    hx, hy, hz = widths
    nx, ny, nz = len(hx), len(hy), len(hz)
    xn = origin[0] + np.r_[0.0, np.cumsum(hx)]
    yn = origin[1] + np.r_[0.0, np.cumsum(hy)]
    zn = origin[2] + np.r_[0.0, np.cumsum(hz)]
    xc, yc, zc = 0.5 * (xn[1:] + xn[:-1]), 0.5 * (yn[1:] + yn[:-1]), 0.5 * (zn[1:] + zn[:-1])
    # median over all nodes of a tensor grid equals the median of the axis coordinates
    s = np.array([np.median(xn), np.median(yn), zn[-1]])

    def alpha(fx, fy, fz, normal):
        r = np.stack(np.broadcast_arrays(fx - s[0], fy - s[1], fz - s[2]))
        rr = np.sqrt((r * r).sum(axis=0))
        with np.errstate(invalid="ignore", divide="ignore"):
            return (r[0] * normal[0] + r[1] * normal[1] + r[2] * normal[2]) / (rr * rr)

    J, K = np.meshgrid(np.arange(ny), np.arange(nz), indexing="ij")
    Jf, Kf = J.ravel(order="F"), K.ravel(order="F")
    I2, K2 = np.meshgrid(np.arange(nx), np.arange(nz), indexing="ij")
    If, Kf2 = I2.ravel(order="F"), K2.ravel(order="F")
    I3, J3 = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    If3, Jf3 = I3.ravel(order="F"), J3.ravel(order="F")
    area_x = hy[Jf] * hz[Kf]
    area_y = hx[If] * hz[Kf2]
    area_z = hx[If3] * hy[Jf3]
    a = {
        "xmin": alpha(xn[0], yc[Jf], zc[Kf], (-1, 0, 0)),
        "xmax": alpha(xn[-1], yc[Jf], zc[Kf], (1, 0, 0)),
        "ymin": alpha(xc[If], yn[0], zc[Kf2], (0, -1, 0)),
        "ymax": alpha(xc[If], yn[-1], zc[Kf2], (0, 1, 0)),
        "zmin": alpha(xc[If3], yc[Jf3], zn[0], (0, 0, -1)),
        "zmax": alpha(xc[If3], yc[Jf3], zn[-1], (0, 0, 1)),
    }
    # discretize boundary-face order: x (2,ny,nz), y (nx,2,nz), z (nx,ny,2), Fortran
    bf_alpha = np.concatenate([
        np.stack([a["xmin"], a["xmax"]], axis=0).ravel(order="F"),
        np.stack([a["ymin"].reshape(nx, nz, order="F"), a["ymax"].reshape(nx, nz, order="F")],
                 axis=1).ravel(order="F"),
        np.concatenate([a["zmin"], a["zmax"]]),
    ])
    if surface_faces is None:
        surf = np.zeros(bf_alpha.size, dtype=bool)
        surf[-nx * ny:] = True
    else:
        surf = np.asarray(surface_faces)
        if surf.dtype != bool:
            mask = np.zeros(bf_alpha.size, dtype=bool)
            mask[surf] = True
            surf = mask
        if surf.size != bf_alpha.size:
            raise ValueError(f"surface_faces must have {bf_alpha.size} entries")
    bf_alpha = np.where(surf, 0.0, bf_alpha)
    # back to the six planes
    nxf = 2 * ny * nz
    nyf = 2 * nx * nz
    ax_ = bf_alpha[:nxf].reshape(2, ny * nz, order="F")
    ay_ = bf_alpha[nxf:nxf + nyf].reshape(nx, 2, nz, order="F")
    az_ = bf_alpha[nxf + nyf:].reshape(2, nx * ny)
    out = {
        "xmin": ax_[0] * area_x / 4.0,
        "xmax": ax_[1] * area_x / 4.0,
        "ymin": ay_[:, 0, :].ravel(order="F") * area_y / 4.0,
        "ymax": ay_[:, 1, :].ravel(order="F") * area_y / 4.0,
        "zmin": az_[0] * area_z / 4.0,
        "zmax": az_[1] * area_z / 4.0,
        "alpha": bf_alpha,
    }
    return out


def edge_mass(widths, sigma=None):
    """Diagonal of the isotropic edge inner product M_e(sigma) = diag(A_v sigma).

    Each edge receives V_c sigma_c / 4 from every cell it borders (discretize edge order:
    x edges (nx, ny+1, nz+1), then y edges, then z edges, each Fortran ordered).
    """
    hx, hy, hz = widths
    nx, ny, nz = len(hx), len(hy), len(hz)
    vol = hx[:, None, None] * hy[None, :, None] * hz[None, None, :]
    s = np.ones((nx, ny, nz)) if sigma is None else np.asarray(sigma, dtype=float).reshape(
        (nx, ny, nz), order="F")
    p = vol * s / 4.0
    px = np.pad(p, ((0, 0), (1, 1), (1, 1)))
    ex = px[:, :-1, :-1] + px[:, 1:, :-1] + px[:, :-1, 1:] + px[:, 1:, 1:]
    py = np.pad(p, ((1, 1), (0, 0), (1, 1)))
    ey = py[:-1, :, :-1] + py[1:, :, :-1] + py[:-1, :, 1:] + py[1:, :, 1:]
    pz = np.pad(p, ((1, 1), (1, 1), (0, 0)))
    ez = pz[:-1, :-1, :] + pz[1:, :-1, :] + pz[:-1, 1:, :] + pz[1:, 1:, :]
    return np.concatenate([ex.ravel(order="F"), ey.ravel(order="F"), ez.ravel(order="F")])


def edge_cell_weights(widths):
    """A_v (n_edges x n_cells, CSR): V_c / 4 for every cell c bordering edge e, so that
    edge_mass(widths, sigma) = A_v @ sigma (the isotropic edge inner product's averaging)."""
    hx, hy, hz = widths
    nx, ny, nz = len(hx), len(hy), len(hz)
    vol = (hx[:, None, None] * hy[None, :, None] * hz[None, None, :]).ravel(order="F") / 4.0

    def cell(i, j, k):
        return i + nx * (j + ny * k)

    rows, cols, vals = [], [], []
    off = 0
    # x edges (i, j, k), i < nx: cells (i, j - a, k - b); y edges: (i - a, j, k - b); z edges:
    # (i - a, j - b, k)
    for axis, shape in ((0, (nx, ny + 1, nz + 1)), (1, (nx + 1, ny, nz + 1)),
                        (2, (nx + 1, ny + 1, nz))):
        I, J, K = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
        e = (I + shape[0] * (J + shape[1] * K)).ravel(order="F") + off
        I, J, K = (X.ravel(order="F") for X in (I, J, K))
        for a in (0, 1):
            for b in (0, 1):
                ci, cj, ck = I.copy(), J.copy(), K.copy()
                if axis == 0:
                    cj, ck = J - a, K - b
                elif axis == 1:
                    ci, ck = I - a, K - b
                else:
                    ci, cj = I - a, J - b
                ok = (ci >= 0) & (ci < nx) & (cj >= 0) & (cj < ny) & (ck >= 0) & (ck < nz)
                c = cell(ci[ok], cj[ok], ck[ok])
                rows.append(e[ok])
                cols.append(c)
                vals.append(vol[c])
        off += int(np.prod(shape))
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(off, nx * ny * nz))


def robin_diagonal(widths, beta, sigma):
    """b_n = sum over boundary faces f containing node n of beta_f sigma(cell of f)."""
    hx, hy, hz = widths
    nx, ny, nz = len(hx), len(hy), len(hz)
    s = np.asarray(sigma, dtype=float).reshape((nx, ny, nz), order="F")
    b = np.zeros((nx + 1, ny + 1, nz + 1))
    bx0 = beta["xmin"].reshape(ny, nz, order="F") * s[0]
    bx1 = beta["xmax"].reshape(ny, nz, order="F") * s[-1]
    for dj in (0, 1):
        for dk in (0, 1):
            b[0, dj:dj + ny, dk:dk + nz] += bx0
            b[-1, dj:dj + ny, dk:dk + nz] += bx1
    by0 = beta["ymin"].reshape(nx, nz, order="F") * s[:, 0, :]
    by1 = beta["ymax"].reshape(nx, nz, order="F") * s[:, -1, :]
    for di in (0, 1):
        for dk in (0, 1):
            b[di:di + nx, 0, dk:dk + nz] += by0
            b[di:di + nx, -1, dk:dk + nz] += by1
    bz0 = beta["zmin"].reshape(nx, ny, order="F") * s[:, :, 0]
    bz1 = beta["zmax"].reshape(nx, ny, order="F") * s[:, :, -1]
    for di in (0, 1):
        for dj in (0, 1):
            b[di:di + nx, dj:dj + ny, 0] += bz0
            b[di:di + nx, dj:dj + ny, -1] += bz1
    return b.ravel(order="F")
