"""DC survey: ordering of sources, receivers and data.

Data are ordered source-major, then receiver, then receiver location
(simpeg/survey.py:416-627).  ``set_geometric_factor`` follows
electromagnetics/static/resistivity/survey.py:168-194 with the geometric factor of
electromagnetics/static/utils/static_utils.py:81-309.
"""

from __future__ import annotations

import uuid

import numpy as np

from . import receivers as Rx
from . import sources as Src


class Survey:
    def __init__(self, source_list, survey_geometry="surface", **kwargs):
        if kwargs.pop("survey_type", None) is not None:
            raise TypeError(
                "Argument 'survey_type' has been removed in SimPEG 0.24.0. Types of sources and"
                "their corresponding receivers are obtained from their respective classes, without "
                "the need to specify the survey type.",
            )
        kwargs.pop("counter", None)
        if kwargs:
            raise TypeError(f"unexpected keyword arguments {sorted(kwargs)}")
        self._uid = uuid.uuid4()
        self.source_list = source_list
        if survey_geometry not in ("surface", "borehole", "general"):
            raise ValueError("survey_geometry must be 'surface', 'borehole' or 'general'")
        self.survey_geometry = survey_geometry

    @property
    def uid(self):
        return self._uid

    @property
    def source_list(self):
        return self._source_list

    @source_list.setter
    def source_list(self, new_list):
        new_list = [] if new_list is None else list(new_list)
        ids = [id(s) for s in new_list]
        if len(set(ids)) != len(ids):
            raise ValueError("source_list must contain unique sources")
        self._source_list = new_list

    @property
    def nSrc(self):
        return len(self.source_list)

    @property
    def vnD(self):
        return np.array([src.nD for src in self.source_list])

    @property
    def nD(self):
        return int(sum(self.vnD))

    def get_slice(self, source, receiver):
        end = 0
        for src in self.source_list:
            for rx in src.receiver_list:
                start, end = end, end + rx.nD
                if src is source and rx is receiver:
                    return slice(start, end)
        raise KeyError(f"Source '{source}' and receiver '{receiver}' pair is not part of the survey.")

    def get_all_slices(self):
        slices = {}
        end = 0
        for src in self.source_list:
            for rx in src.receiver_list:
                start, end = end, end + rx.nD
                slices[(src, rx)] = slice(start, end)
        return slices

    def __repr__(self):
        return f"{self.__class__.__name__}(#sources: {self.nSrc}; #data: {self.nD})"

    # electrode geometry -------------------------------------------------------------------
    def _abmn_locations(self):
        """Per-datum A, B, M, N electrode locations (survey.py:196-237): a pole source repeats A
        as B, a pole receiver repeats M as N."""
        a, b, m, n = [], [], [], []
        for src in self.source_list:
            for rx in src.receiver_list:
                k = rx.nD
                if isinstance(src, Src.Pole):
                    a.append(np.repeat(src.location.reshape(1, -1), k, axis=0))
                    b.append(np.repeat(src.location.reshape(1, -1), k, axis=0))
                elif isinstance(src, Src.Dipole):
                    a.append(np.repeat(src.location[0].reshape(1, -1), k, axis=0))
                    b.append(np.repeat(src.location[1].reshape(1, -1), k, axis=0))
                else:
                    raise NotImplementedError("A/B locations are undefined for multipole sources")
                if isinstance(rx, Rx.Pole):
                    m.append(rx.locations)
                    n.append(rx.locations)
                else:
                    m.append(rx.locations[0])
                    n.append(rx.locations[1])
        return np.vstack(a), np.vstack(b), np.vstack(m), np.vstack(n)

    @property
    def locations_a(self):
        return self._abmn_locations()[0]

    @property
    def locations_b(self):
        return self._abmn_locations()[1]

    @property
    def locations_m(self):
        return self._abmn_locations()[2]

    @property
    def locations_n(self):
        return self._abmn_locations()[3]

    @property
    def unique_electrode_locations(self):
        return np.unique(np.vstack(self._abmn_locations()), axis=0)

    def _abmn(self):
        a, b, m, n = [], [], [], []
        for src in self.source_list:
            if isinstance(src, Src.Dipole):
                aa, bb = src.location[0], src.location[1]
            elif isinstance(src, Src.Pole):
                aa, bb = src.location[0], np.inf * np.ones_like(src.location[0])
            else:
                raise NotImplementedError("A_B locations for undefined for multipole sources.")
            for rx in src.receiver_list:
                if isinstance(rx, Rx.Dipole):
                    M, N = rx.locations[0], rx.locations[1]
                else:
                    M, N = rx.locations, -np.inf * np.ones_like(rx.locations)
                k = M.shape[0]
                a.append(np.tile(aa, (k, 1)))
                b.append(np.tile(bb, (k, 1)))
                m.append(M)
                n.append(N)
        return np.vstack(a), np.vstack(b), np.vstack(m), np.vstack(n)

    def set_geometric_factor(self, space_type="halfspace"):
        """Geometric factor (1/AM - 1/BM - 1/AN + 1/BN) / (C pi), C = 2 (half space) or 4."""
        st = space_type.lower().replace("-", "").replace("_", "").replace(" ", "")
        if st == "halfspace":
            fact = 2.0
        elif st == "wholespace":
            fact = 4.0
        else:
            raise TypeError("'space_type must be 'whole space' | 'half space'")
        A, B, M, N = self._abmn()
        with np.errstate(divide="ignore", invalid="ignore"):
            AM = np.linalg.norm(A - M, axis=1)
            BM = np.linalg.norm(B - M, axis=1)
            AN = np.linalg.norm(A - N, axis=1)
            BN = np.linalg.norm(B - N, axis=1)
            G = (1 / AM - 1 / BM - 1 / AN + 1 / BN) / (fact * np.pi)
        slices = self.get_all_slices()
        for src in self.source_list:
            for rx in src.receiver_list:
                if rx.data_type == "apparent_resistivity":
                    rx._geometric_factor[src] = G[slices[src, rx]]
        return G
