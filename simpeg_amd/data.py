"""Observed / synthetic data containers (simpeg/data.py: ``Data``, ``SyntheticData``).

standard_deviation = sqrt((relative_error |dobs|)^2 + noise_floor^2) (data.py:223-262).
"""

from __future__ import annotations

import numpy as np


class Data:
    def __init__(self, survey, dobs=None, relative_error=None, noise_floor=None,
                 standard_deviation=None):
        self.survey = survey
        if dobs is None:
            dobs = np.full(survey.nD, np.nan)
        dobs = np.asarray(dobs, dtype=float).ravel()
        if dobs.size != survey.nD:
            raise ValueError(f"dobs must have {survey.nD} entries, got {dobs.size}")
        self.dobs = dobs
        self.relative_error = relative_error
        self.noise_floor = noise_floor
        if standard_deviation is not None:
            if relative_error is not None or noise_floor is not None:
                raise TypeError("give standard_deviation or relative_error/noise_floor, not both")
            self.noise_floor = standard_deviation
            self.relative_error = None

    @property
    def nD(self):
        return self.dobs.size

    @property
    def standard_deviation(self):
        if self.relative_error is None and self.noise_floor is None:
            raise TypeError("the standard deviation needs relative_error and/or noise_floor")
        uncert = np.zeros(self.nD)
        if self.relative_error is not None:
            uncert += np.asarray(self.relative_error * np.abs(self.dobs)) ** 2
        if self.noise_floor is not None:
            uncert += np.asarray(self.noise_floor) ** 2
        return np.sqrt(uncert)


class SyntheticData(Data):
    def __init__(self, survey, dobs=None, dclean=None, relative_error=None, noise_floor=None):
        super().__init__(survey, dobs=dobs, relative_error=relative_error, noise_floor=noise_floor)
        self.dclean = None if dclean is None else np.asarray(dclean, dtype=float)
