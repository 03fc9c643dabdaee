"""Nodal induced polarisation on the GPU (electromagnetics/static/induced_polarization/
simulation.py:15-139, Simulation3DNodal at :160).

IP data are linear in the chargeability eta: d = s * J_dc(sigma) eta with the DC sensitivity
taken through sigmaDeriv = -diag(sigma) etaDeriv, and s = 1 / (DC voltage) for
``apparent_chargeability`` receivers, 1 otherwise.  The DC fields of the fixed conductivity are
solved once and reused; Jvec / Jtvec are the DC ones (batched solves + contraction kernel)
scaled by s.
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from .maps import sdiag
from .resistivity.simulation import Simulation3DNodal as _DCNodal


class Simulation3DNodal(_DCNodal):
    def __init__(self, mesh, survey=None, sigma=None, rho=None, eta=None, etaMap=None, Ainv=None,
                 _f=None, **kwargs):
        kwargs.pop("sigmaMap", None)
        kwargs.pop("rhoMap", None)
        super().__init__(mesh, survey=survey, sigma=sigma, rho=rho, **kwargs)
        self.etaMap = etaMap
        self._eta_fixed = None if eta is None else np.asarray(eta, dtype=float).ravel()
        self._f = _f
        self._pred = None
        self._scale_cache = None

    # the model is the chargeability; sigma is fixed
    @property
    def _active_map(self):
        return self.etaMap

    # the reference IP class reports an IdentityMap as its sigmaMap, so its getADeriv keeps the
    # Robin term (induced_polarization/simulation.py:19-25; resistivity/simulation.py:566-570)
    @property
    def _deriv_includes_bc(self):
        return True

    @property
    def sigma(self):
        if self._sigma_fixed is not None:
            return self._sigma_fixed
        if self._rho_fixed is not None:
            return 1.0 / self._rho_fixed
        raise AttributeError("IP simulations need a fixed sigma or rho")

    @property
    def eta(self):
        if self.etaMap is not None:
            return np.asarray(self.etaMap * self.model, dtype=float).ravel()
        return self._eta_fixed

    @property
    def etaDeriv(self):
        if self.etaMap is None:
            raise AttributeError("etaDeriv needs etaMap")
        return sp.csr_matrix(self.etaMap.deriv(self.model))

    @property
    def sigmaDeriv(self):
        return sp.csr_matrix(sdiag(-self.sigma) @ self.etaDeriv)

    @property
    def rhoDeriv(self):
        return sp.csr_matrix(sdiag(self.rho) @ self.etaDeriv)

    def _ensure_sigma(self):
        # sigma does not depend on the model: upload it once
        if self._sigma_on_gpu is None:
            self.engine.set_sigma(self.sigma)
            self._sigma_on_gpu = True

    @property
    def model(self):
        return self._model

    @model.setter
    def model(self, value):
        self._assign_model(value, copy=True)

    def _assign_model(self, value, copy):
        if value is None:
            self._model = None
            return
        value = np.asarray(value, dtype=float).ravel()
        mp = self.etaMap
        if mp is not None and mp.nP != "*" and value.size != mp.nP:
            raise ValueError(f"Model is the wrong size: {value.size}, expected {mp.nP} parameters")
        if self._model is None or self._model.shape != value.shape or not np.allclose(
                self._model, value):
            self._Jmatrix = None
            self._gtgdiag = None
            self._model_version += 1
        self._model = value.copy() if copy else value

    def _dc_fields(self):
        if self._f is None:
            self._f = _DCNodal.fields(self, None)
        return self._f

    @property
    def _scale(self):
        if self._scale_cache is None:
            survey = self.survey
            scale = np.ones(survey.nD)
            slices = survey.get_all_slices()
            need = [(src, rx) for src in survey.source_list for rx in src.receiver_list
                    if rx.data_type == "apparent_chargeability"]
            if need:
                dc = _DCNodal.dpred(self, None, f=self._dc_fields())
                for src, rx in need:
                    sl = slices[src, rx]
                    scale[sl] = 1.0 / dc[sl]
            self._scale_cache = scale
        return self._scale_cache

    def fields(self, m=None, calcJ=True):
        f = self._dc_fields()
        if m is not None:
            self._pred = self.forward(m, f=f)
        return f

    def forward(self, m, f=None):
        return np.asarray(self.Jvec(m, m, f=f))

    def dpred(self, m=None, f=None):
        if m is None:
            m = self.model
        if f is None or self._pred is None:
            self.fields(m)
        return self._pred

    def Jvec(self, m, v, f=None):
        return self._scale * _DCNodal.Jvec(self, m, v, f=self._dc_fields() if f is None else f)

    def Jtvec(self, m, v, f=None):
        return _DCNodal.Jtvec(self, m, np.asarray(v, dtype=float) * self._scale,
                              f=self._dc_fields() if f is None else f)

    def getJtJdiag(self, m, W=None, f=None):
        if self._gtgdiag is None:
            J = self.getJ(m, f=self._dc_fields() if f is None else f)
            w = self._scale ** 2 if W is None else (self._scale * np.asarray(W.diagonal())) ** 2
            self._gtgdiag = np.einsum("i,ij,ij->j", w, J, J)
        return self._gtgdiag
