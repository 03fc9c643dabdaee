"""Survey miniaturisation (behaviour of electromagnetics/static/resistivity/utils.py:75-128 and
simulation.py:296-318).

A datum of a dipole or pole survey is V(A,M) - V(A,N) - V(B,M) + V(B,N), where V(X,Y) is the
potential at electrode Y of a unit pole source at electrode X; a pole source has no B term and a
pole receiver no N term.  A is symmetric, so V(X,Y) = V(Y,X): every term is named by an
unordered electrode pair.  The survey is reduced to its distinct pairs, solved as pole sources
with pole receivers, and recombined by a sparse (nD x n_pairs) matrix T of +-1 entries:
d = T d_pairs, and the adjoint is T^T v.
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from . import receivers, sources
from .survey import Survey


def _mini_pole_pole(survey, verbose=False):
    """(T, mini_survey) for ``survey``.

    Electrodes are numbered by their sorted unique coordinates.  A pair {i, j} is keyed as
    min * n + max; the mini survey has one pole source per distinct smaller index, with pole
    receivers at its partners in increasing order, so row r of the mini data is pair key r in
    sorted order.
    """
    loc = np.concatenate([survey.locations_a, survey.locations_b, survey.locations_m,
                          survey.locations_n])
    elecs, code = np.unique(loc, axis=0, return_inverse=True)
    code = np.asarray(code).ravel()
    nd = survey.nD
    a, b, m, n = code[:nd], code[nd:2 * nd], code[2 * nd:3 * nd], code[3 * nd:]
    has_b = a != b
    has_n = m != n
    ne = len(elecs)
    datum = np.arange(nd)
    # (datum, source electrode, receiver electrode, sign) of the four potential terms
    parts = [(datum, a, m, np.ones(nd)),
             (datum[has_n], a[has_n], n[has_n], -np.ones(has_n.sum())),
             (datum[has_b], b[has_b], m[has_b], -np.ones(has_b.sum())),
             (datum[has_b & has_n], b[has_b & has_n], n[has_b & has_n],
              np.ones((has_b & has_n).sum()))]
    rows = np.concatenate([p[0] for p in parts])
    x = np.concatenate([p[1] for p in parts]).astype(np.int64)
    y = np.concatenate([p[2] for p in parts]).astype(np.int64)
    sign = np.concatenate([p[3] for p in parts])
    keys = np.minimum(x, y) * ne + np.maximum(x, y)
    ukeys, col = np.unique(keys, return_inverse=True)
    T = sp.csr_matrix((sign, (rows, np.asarray(col).ravel())), shape=(nd, len(ukeys)))
    if verbose:
        print(f"There are {len(ukeys)} unique pole-pole combinations.")
    lo, hi = np.divmod(ukeys, ne)
    cut = np.flatnonzero(np.diff(lo)) + 1
    src_list = [sources.Pole([receivers.Pole(elecs[hi[g]])], elecs[lo[g[0]]])
                for g in np.split(np.arange(len(ukeys)), cut)]
    return T, Survey(src_list)


def mini_survey_data(T, d_mini):
    """Full-survey data from the pole-pole data (simulation.py:296-304)."""
    return np.asarray(T @ np.asarray(d_mini, dtype=float))


def mini_survey_dataT(T, v):
    """Adjoint of ``mini_survey_data`` (simulation.py:306-318)."""
    return np.asarray(T.T @ np.asarray(v, dtype=float))
