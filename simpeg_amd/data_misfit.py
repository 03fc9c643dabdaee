"""``L2DataMisfit`` with the reference's conventions (simpeg/data_misfit.py:149-343).

phi(m) = ||W (dpred(m) - dobs)||^2 (no 1/2), deriv = 2 J^T W^T W r,
deriv2(m, v) = 2 J^T W^T W J v.  W defaults to diag(1 / standard_deviation).
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp


class L2DataMisfit:
    def __init__(self, data, simulation, W=None):
        self.data = data
        self.simulation = simulation
        self._W = W

    @property
    def W(self):
        if self._W is None:
            self._W = sp.diags(1.0 / self.data.standard_deviation, format="csr")
        return self._W

    @W.setter
    def W(self, value):
        if value is not None and not sp.issparse(value):
            value = sp.diags(np.asarray(value, dtype=float), format="csr")
        self._W = value

    @property
    def nD(self):
        return self.data.nD

    def residual(self, m, f=None):
        return self.simulation.residual(m, self.data.dobs, f=f)

    def __call__(self, m, f=None):
        r = self.W @ self.residual(m, f=f)
        return float(r @ r)

    def deriv(self, m, f=None):
        if f is None:
            f = self.simulation.fields(m)
        return 2.0 * self.simulation.Jtvec(m, self.W.T @ (self.W @ self.residual(m, f=f)), f=f)

    def deriv2(self, m, v, f=None):
        if f is None:
            f = self.simulation.fields(m)
        return 2.0 * self.simulation.Jtvec(
            m, self.W.T @ (self.W @ self.simulation.Jvec(m, v, f=f)), f=f)
