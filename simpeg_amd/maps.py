"""Model -> physical property maps used by the nodal DC simulation.

Same call surface as ``simpeg.maps``: ``map * m`` (or ``map @ m``) evaluates the map,
``map.deriv(m)`` returns the (sparse) derivative and ``map.deriv(m, v)`` its product
with ``v``; ``nP`` is the number of model parameters.  Reference: ``IdentityMap``
(maps/_base.py:26-256), ``ExpMap`` (maps/_property_maps.py:18-112),
``ReciprocalMap`` (maps/_property_maps.py), ``ComboMap`` (maps/_base.py:357-548).
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp


def sdiag(d):
    """Diagonal CSR matrix of the vector d, built directly (scipy.sparse.diags goes through the
    DIA and COO formats, ~80 ms at a million entries)."""
    d = np.asarray(d, dtype=float).ravel()
    n = d.size
    idx = np.arange(n + 1, dtype=np.int64 if n >= 2**31 - 1 else np.int32)
    return sp.csr_matrix((d, idx[:n], idx), shape=(n, n))


class IdentityMap:
    def __init__(self, mesh=None, nP=None, **kwargs):
        if nP is None and mesh is not None:
            nP = mesh.n_cells if hasattr(mesh, "n_cells") else mesh.nC
        self._nP = nP

    @property
    def nP(self):
        return "*" if self._nP is None else int(self._nP)

    @property
    def shape(self):
        return (self.nP, self.nP)

    @property
    def is_linear(self):
        return True

    def _transform(self, m):
        return np.asarray(m, dtype=float).ravel()

    def deriv(self, m, v=None):
        d = sdiag(np.ones(np.size(m)))
        return d if v is None else d @ v

    def __mul__(self, val):
        if isinstance(val, (IdentityMap, ComboMap)):
            return ComboMap([self, val])
        return self._transform(val)

    __matmul__ = __mul__

    def inverse(self, D):
        return np.asarray(D, dtype=float).ravel()

    # Elementwise maps also evaluate on device tensors (torch), so a model that lives on the GPU
    # never goes through host memory: the value and the diagonal of the derivative.
    def torch_transform(self, m):
        return m

    def torch_deriv_diag(self, m):
        return m.new_ones(m.shape)


class ExpMap(IdentityMap):
    """sigma = exp(m); derivative diag(exp(m))."""

    def _transform(self, m):
        return np.exp(np.asarray(m, dtype=float).ravel())

    def inverse(self, D):
        return np.log(np.asarray(D, dtype=float).ravel())

    def deriv(self, m, v=None):
        d = sdiag(np.exp(np.asarray(m, dtype=float).ravel()))
        return d if v is None else d @ v

    def torch_transform(self, m):
        return m.exp()

    def torch_deriv_diag(self, m):
        return m.exp()

    @property
    def is_linear(self):
        return False


class ReciprocalMap(IdentityMap):
    """sigma = 1/m; derivative diag(-1/m^2)."""

    def _transform(self, m):
        return 1.0 / np.asarray(m, dtype=float).ravel()

    def inverse(self, D):
        return 1.0 / np.asarray(D, dtype=float).ravel()

    def deriv(self, m, v=None):
        m = np.asarray(m, dtype=float).ravel()
        d = sdiag(-1.0 / m ** 2)
        return d if v is None else d @ v

    def torch_transform(self, m):
        return 1.0 / m

    def torch_deriv_diag(self, m):
        return -1.0 / (m * m)

    @property
    def is_linear(self):
        return False


class ComboMap(IdentityMap):
    """Composition maps[0] * maps[1] * ... * m (right-most applied first)."""

    def __init__(self, maps, **kwargs):
        super().__init__(nP=None)
        self.maps = []
        for m in maps:
            if isinstance(m, ComboMap):
                self.maps.extend(m.maps)
            else:
                self.maps.append(m)

    @property
    def nP(self):
        return self.maps[-1].nP

    torch_transform = None
    torch_deriv_diag = None

    def _transform(self, m):
        for mp in reversed(self.maps):
            m = mp * m
        return m

    def deriv(self, m, v=None):
        deriv = None
        mi = m
        for mp in reversed(self.maps):
            d = mp.deriv(mi)
            deriv = d if deriv is None else d @ deriv
            mi = mp * mi
        return deriv if v is None else deriv @ v

    @property
    def is_linear(self):
        return all(getattr(m, "is_linear", False) for m in self.maps)


class Projection(IdentityMap):
    """Select entries ``index`` of an nP-long model (maps/_base.py ``Projection``)."""

    def __init__(self, nP, index, **kwargs):
        super().__init__(nP=nP)
        self.index = np.atleast_1d(np.arange(nP)[index] if isinstance(index, slice) else index)
        # each model entry selected at most once: P^T v is a scatter into a zero vector
        self.unique_index = bool(np.unique(self.index).size == self.index.size)
        # a contiguous ascending range is copied as one block instead of gathered
        i = self.index
        self._range = (slice(int(i[0]), int(i[0]) + i.size)
                       if i.size and np.array_equal(i, np.arange(i[0], i[0] + i.size)) else None)
        self._P = sp.csr_matrix((np.ones(self.index.size), (np.arange(self.index.size), self.index)),
                                shape=(self.index.size, nP))

    @property
    def shape(self):
        return (self.index.size, self.nP)

    torch_transform = None
    torch_deriv_diag = None

    @property
    def is_linear(self):
        return True

    def _transform(self, m):
        m = np.asarray(m, dtype=float).ravel()
        return m[self._range].copy() if self._range is not None else m[self.index]

    def deriv(self, m, v=None):
        return self._P if v is None else self._P @ v
