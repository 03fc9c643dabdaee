"""A pymatsolver-shaped GPU solver for the assembled nodal DC operator.

The reference lets a user pick the linear solver: `Simulation3DNodal(..., solver=S,
solver_opts={...})` builds `Ainv = S(A, **solver_opts)` from the assembled CSR operator and applies
it as `Ainv * rhs` for one or many right-hand sides, then frees it with `Ainv.clean()`
(resistivity/simulation.py:105-111, 181, 245; the solver contract is
base/pde_simulation.py:493-511).  `NodalGPUSolver` keeps that calling convention: it reads the
7-point tensor-mesh structure back out of A (edge conductances from the sub-diagonals, the
Robin term from the rest of the diagonal; a Neumann operator is recognised by its identity row 0)
and solves with the GPU multigrid-preconditioned CG.  It needs the mesh, given in solver_opts.
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from .engine import NodalEngine
from .mesh import mesh_widths


def nodal_coefficients(A, mesh):
    """(cx, cy, cz, bd, neumann) of an assembled nodal operator A of a 3D tensor mesh.

    cx/cy/cz[n]: conductance of the +x/+y/+z edge of node n (0 where there is none), read from
    the lower diagonals (row n + s holds -w of edge (n, n + s); row 0 is never needed, so a
    Neumann operator's replaced row 0 does not matter); bd: the rest of the diagonal.  Raises
    ValueError unless A is exactly the symmetric 7-point stencil these describe (apart from the
    identity row 0 of the Neumann form)."""
    A = sp.csr_matrix(A, dtype=np.float64)
    widths, _ = mesh_widths(mesh)
    n1x, n1y, n1z = (len(h) + 1 for h in widths)
    nN = n1x * n1y * n1z
    if A.shape != (nN, nN):
        raise ValueError(f"A has shape {A.shape}, the mesh has {nN} nodes")
    plane = n1x * n1y
    row0 = A.getrow(0)
    neumann = bool(row0.nnz == 1 and row0.indices[0] == 0 and row0.data[0] == 1.0 and
                   A[1, 0] != 0.0)
    cx = np.zeros(nN)
    cy = np.zeros(nN)
    cz = np.zeros(nN)
    cx[: nN - 1] = -A.diagonal(-1)
    cy[: nN - n1x] = -A.diagonal(-n1x)
    cz[: nN - plane] = -A.diagonal(-plane)
    idx = np.arange(nN)
    cx[idx % n1x == n1x - 1] = 0.0
    cy[(idx // n1x) % n1y == n1y - 1] = 0.0
    cz[idx // plane == n1z - 1] = 0.0
    wsum = cx + cy + cz
    wsum[1:] += cx[:-1]
    wsum[n1x:] += cy[:-n1x]
    wsum[plane:] += cz[:-plane]
    bd = A.diagonal() - wsum
    if neumann:
        bd[0] = 0.0
    B = sp.diags([A.diagonal(), -cx[:-1], -cx[:-1], -cy[:-n1x], -cy[:-n1x], -cz[:-plane],
                  -cz[:-plane]], [0, 1, -1, n1x, -n1x, plane, -plane], shape=(nN, nN),
                 format="csr")
    if neumann:
        B = B.tolil()
        B[0, :] = 0.0
        B[0, 0] = 1.0
        B = B.tocsr()
    diff = abs(A - B)
    if diff.nnz and diff.max() > 1e-12 * abs(A).max():
        raise ValueError("A is not the 7-point nodal operator of this tensor mesh")
    if (cx < 0).any() or (cy < 0).any() or (cz < 0).any():
        raise ValueError("A has positive off-diagonal entries (negative edge conductances)")
    return cx, cy, cz, bd, neumann


class NodalGPUSolver:
    """Ainv = NodalGPUSolver(A, mesh=mesh); x = Ainv * b  (b: (nN,) or (nN, k))."""

    def __init__(self, A, mesh=None, rtol=1e-10, precision="mixed", max_iter=1000, device=None,
                 gpu_params=None, **kwargs):
        if mesh is None:
            raise TypeError("NodalGPUSolver needs the tensor mesh: pass solver_opts={'mesh': mesh}")
        cx, cy, cz, bd, neumann = nodal_coefficients(A, mesh)
        widths, origin = mesh_widths(mesh)
        self.neumann = neumann
        self.rtol = float(rtol)
        self.precision = precision
        self.max_iter = int(max_iter)
        self.engine = NodalEngine(widths, origin, bc_type="Neumann" if neumann else "Robin",
                                  device=device, params=gpu_params)
        self.engine.set_coefficients(cx, cy, cz, bd)
        self.last_solve = None

    def solve(self, rhs):
        b = np.asarray(rhs, dtype=np.float64)
        vec = b.ndim == 1
        B = b[:, None] if vec else b
        if B.ndim != 2 or B.shape[0] != self.engine.nN:
            raise ValueError(f"rhs must have shape ({self.engine.nN},) or ({self.engine.nN}, k), "
                             f"got {b.shape}")
        if B.shape[1] == 0:
            return np.zeros_like(b)
        eng = self.engine
        q = eng.to_block(B)
        u, res = eng.solve(q, B.shape[1], mode=self.precision, tol=self.rtol,
                           max_iter=self.max_iter)
        self.last_solve = res
        x = eng.from_block(u, B.shape[1])
        return x[:, 0] if vec else x

    def __mul__(self, rhs):
        return self.solve(rhs)

    __matmul__ = __mul__

    def clean(self):
        """Release the GPU handle (the reference calls Ainv.clean() before refactorising)."""
        if getattr(self, "engine", None) is not None:
            self.engine.close()
            self.engine = None
