"""DC sources (``Dipole``, ``Pole``, ``Multipole``) with the reference's constructor
semantics (electromagnetics/static/resistivity/sources.py:6-270, survey.py:282-415).

``BaseSrc.eval(sim)`` returns the nodal right-hand side
q = sum_k I_k * trilinear_node_weights(location_k) (sources.py:64-88).
"""

from __future__ import annotations

import uuid

import numpy as np

from ..mesh import mesh_widths, node_interpolation


def _as_2d(name, val):
    arr = np.atleast_2d(np.asarray(val, dtype=float))
    if arr.ndim != 2:
        raise ValueError(f"{name} must be a (n, dim) array")
    return arr


class BaseSrc:
    _fields_per_source = 1

    def __init__(self, receiver_list=None, location=None, current=1.0, **kwargs):
        if kwargs:
            raise TypeError(f"unexpected keyword arguments {sorted(kwargs)}")
        self._uid = uuid.uuid4()
        self.receiver_list = [] if receiver_list is None else receiver_list
        self.location = location
        self.current = current

    @property
    def uid(self):
        return self._uid

    @property
    def receiver_list(self):
        return self._receiver_list

    @receiver_list.setter
    def receiver_list(self, value):
        value = list(value)
        ids = [id(r) for r in value]
        if len(set(ids)) != len(ids):
            raise ValueError("receiver_list must contain unique receivers")
        self._receiver_list = value

    @property
    def location(self):
        return self._location

    @location.setter
    def location(self, other):
        self._location = _as_2d("location", other)

    @property
    def current(self):
        return self._current

    @current.setter
    def current(self, other):
        other = np.atleast_1d(np.asarray(other, dtype=float)).ravel()
        if len(other) > 1 and len(other) != self.location.shape[0]:
            raise ValueError(
                "Current must be constant or equal to the number of specified source locations."
                f" saw {len(other)} current sources and {self.location.shape[0]} locations."
            )
        self._current = other

    @property
    def nD(self):
        return int(sum(rx.nD for rx in self.receiver_list))

    @property
    def vnD(self):
        return np.array([rx.nD for rx in self.receiver_list])

    def electrodes(self):
        """(locations, currents) of the current electrodes, currents broadcast."""
        cur = self.current
        if cur.size == 1:
            cur = np.full(self.location.shape[0], cur[0])
        return self.location, cur

    def eval(self, sim):  # noqa: A003
        """Nodal right-hand side of this source on ``sim.mesh`` (dense, n_nodes)."""
        widths, origin = mesh_widths(sim.mesh)
        locs, cur = self.electrodes()
        idx, w = node_interpolation(widths, origin, locs)
        q = np.zeros(int(np.prod([len(h) + 1 for h in widths])))
        np.add.at(q, idx.ravel(), (cur[:, None] * w).ravel())
        return q

    def evalDeriv(self, sim):
        return 0.0


class Multipole(BaseSrc):
    @property
    def location_a(self):
        return self.location

    @property
    def location_b(self):
        return np.full_like(self.location, np.nan)


class Dipole(BaseSrc):
    """Dipole source: currents [+I, -I] at [location_a, location_b]."""

    def __init__(self, receiver_list, location_a=None, location_b=None, location=None, current=1.0):
        if location is None and location_a is None and location_b is None:
            raise TypeError(
                "Found 'location', 'location_a' and 'location_b' as None. "
                "Please specify 'location', or 'location_a' and 'location_b' "
                "when defining a dipole source."
            )
        if location is not None and (location_a is not None or location_b is not None):
            raise TypeError(
                "Found 'location_a' and/or 'location_b' as not None values. "
                "When passing a not None value for 'location', 'location_a' and "
                "'location_b' should be set to None."
            )
        if location is None:
            if location_a is None:
                raise TypeError(
                    "Invalid 'location_a' set to None. When 'location' is None, "
                    "both 'location_a' and 'location_b' should be set to "
                    "a value different than None."
                )
            if location_b is None:
                raise TypeError(
                    "Invalid 'location_b' set to None. When 'location' is None, "
                    "both 'location_a' and 'location_b' should be set to "
                    "a value different than None."
                )
            location = [location_a, location_b]
        if len(location) != 2:
            raise ValueError(
                "location must be a list or tuple of length 2: "
                "[location_a, location_b]. The input location has "
                f"length {len(location)}"
            )
        super().__init__(receiver_list=receiver_list, location=location, current=[current, -current])

    def __repr__(self):
        return f"{self.__class__.__name__}(a: {self.location_a}; b: {self.location_b})"

    @property
    def location_a(self):
        return self.location[0]

    @property
    def location_b(self):
        return self.location[1]


class Pole(BaseSrc):
    """Pole source: one electrode carrying the current."""

    def __init__(self, receiver_list, location=None, **kwargs):
        super().__init__(receiver_list=receiver_list, location=location, **kwargs)
        if len(self.location) != 1:
            raise ValueError(f"Pole sources only have a single location, not {len(self.location)}")

    @property
    def location_a(self):
        return self.location[0]

    @property
    def location_b(self):
        return np.full_like(self.location[0], np.nan)
