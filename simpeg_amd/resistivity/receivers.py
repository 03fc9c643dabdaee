"""DC receivers (``Dipole``, ``Pole``) with the reference's constructor and data-type
semantics (electromagnetics/static/resistivity/receivers.py:10-498).

The projection of a receiver is P = interp(M) - interp(N) (dipole) or interp(M) (pole),
scaled by 1/geometric factor for apparent-resistivity data (receivers.py:156-244).
"""

from __future__ import annotations

import uuid

import numpy as np

_DATA_TYPES = {
    "volt": ("volt", "potential", "potentials", "v", "voltage", "voltages"),
    "apparent_resistivity": (
        "apparent_resistivity", "apparent resistivity", "appresistivity",
        "apparentresistivity", "apparent-resistivity", "appres",
    ),
    "apparent_chargeability": (
        "apparent_chargeability", "apparent chargeability", "appchargeability",
        "apparentchargeability", "apparent-chargeability",
    ),
}


def _validate_data_type(var):
    if not isinstance(var, str):
        raise TypeError("data_type must be a str")
    low = var.lower()
    for key, names in _DATA_TYPES.items():
        if low in names:
            return key
    raise ValueError(f"data_type must be one of {list(_DATA_TYPES)}, got {var!r}")


class BaseRx:
    def __init__(self, locations, data_type="volt", orientation=None, projField="phi",
                 storeProjections=False, **kwargs):
        kwargs.pop("threshold", None)
        if kwargs:
            raise TypeError(f"unexpected keyword arguments {sorted(kwargs)}")
        self._uid = uuid.uuid4()
        self.locations = locations
        self.orientation = orientation
        self.data_type = data_type
        self.projField = projField
        self.storeProjections = storeProjections
        self._geometric_factor = {}

    @property
    def uid(self):
        return self._uid

    @property
    def orientation(self):
        return self._orientation

    @orientation.setter
    def orientation(self, var):
        if var is not None and var not in ("x", "y", "z"):
            raise ValueError("orientation must be None, 'x', 'y' or 'z'")
        self._orientation = var

    @property
    def projField(self):
        return self._projField

    @projField.setter
    def projField(self, var):
        if var not in ("phi", "e", "j"):
            raise ValueError("projField must be 'phi', 'e' or 'j'")
        if var != "phi":
            raise NotImplementedError("the GPU nodal simulation projects phi only")
        self._projField = var

    @property
    def data_type(self):
        return self._data_type

    @data_type.setter
    def data_type(self, var):
        self._data_type = _validate_data_type(var)

    @property
    def geometric_factor(self):
        return self._geometric_factor

    def _electrode_sets(self):
        """[(locations, sign)] of the potential electrodes."""
        raise NotImplementedError


class Dipole(BaseRx):
    """Dipole receiver: v = phi(M) - phi(N)."""

    def __init__(self, locations_m=None, locations_n=None, locations=None, **kwargs):
        if locations_m is not None or locations_n is not None:
            if locations_n is None or locations_m is None:
                raise ValueError("For a dipole receiver both locations_m and locations_n must be set")
            if locations is not None:
                raise ValueError(
                    "Cannot set both locations and locations_m, locations_n. "
                    "Please provide either locations=(locations_m, locations_n) "
                    "or both locations_m=locations_m, locations_n=locations_n"
                )
            locations = [locations_m, locations_n]
        if locations is None:
            raise AttributeError(
                "Receiver cannot be instantiated without assigning 'locations'."
                "Please provide either locations=(locations_m, locations_n) "
                "or both locations_m=locations_m, locations_n=locations_n"
            )
        super().__init__(locations=locations, **kwargs)

    @property
    def locations(self):
        return self._locations

    @locations.setter
    def locations(self, locs):
        if len(locs) != 2:
            raise ValueError(
                "locations must be a list or tuple of length 2: "
                "[locations_m, locations_n]. The input locations has "
                f"length {len(locs)}"
            )
        m = np.atleast_2d(np.asarray(locs[0], dtype=float))
        n = np.atleast_2d(np.asarray(locs[1], dtype=float))
        if m.shape != n.shape:
            raise ValueError(
                f"locations_m (shape: {m.shape}) and locations_n (shape: {n.shape}) need to be the same size"
            )
        self._locations = [m, n]

    @property
    def locations_m(self):
        return self.locations[0]

    @property
    def locations_n(self):
        return self.locations[1]

    @property
    def nD(self):
        return self.locations[0].shape[0]

    def _electrode_sets(self):
        return [(self.locations[0], 1.0), (self.locations[1], -1.0)]


class Pole(BaseRx):
    """Pole receiver: v = phi(M)."""

    @property
    def locations(self):
        return self._locations

    @locations.setter
    def locations(self, locs):
        self._locations = np.atleast_2d(np.asarray(locs, dtype=float))

    @property
    def nD(self):
        return self.locations.shape[0]

    def _electrode_sets(self):
        return [(self.locations, 1.0)]
