"""Survey miniaturisation (electromagnetics/static/resistivity/utils.py:75-128).

Every datum of a dipole/pole survey is a combination AM - AN - BM + BN of pole-pole
potentials.  By reciprocity (A is symmetric) a pole-pole pair is the same whichever electrode
is the source, so the survey reduces to the unique (sorted) electrode pairs: one pole source
per unique first electrode with pole receivers at its partners.  ``dipoles`` flags which data
have a B (dipole_tx) and an N (dipole_rx) electrode; ``invs`` maps every datum's AM, AN, BM, BN
term to its row in the mini survey's data vector.
"""

from __future__ import annotations

import numpy as np

from . import receivers, sources
from .survey import Survey


def _mini_pole_pole(survey, verbose=False):
    A = survey.locations_a
    B = survey.locations_b
    M = survey.locations_m
    N = survey.locations_n
    elecs, inverse = np.unique(np.r_[A, B, M, N], axis=0, return_inverse=True)
    inverse = np.asarray(inverse).ravel()
    inv_A, inv_B, inv_M, inv_N = inverse.reshape(4, -1)
    dipole_tx = inv_A != inv_B
    dipole_rx = inv_M != inv_N

    AM = np.sort(np.c_[inv_A, inv_M])
    AN = np.sort(np.c_[inv_A[dipole_rx], inv_N[dipole_rx]])
    BM = np.sort(np.c_[inv_B[dipole_tx], inv_M[dipole_tx]])
    BN = np.sort(np.c_[inv_B[dipole_tx & dipole_rx], inv_N[dipole_tx & dipole_rx]])
    pairs, pair_inv = np.unique(np.r_[AM, AN, BM, BN], axis=0, return_inverse=True)
    pair_inv = np.asarray(pair_inv).ravel()
    inv_AM, rest = pair_inv[: len(AM)], pair_inv[len(AM):]
    inv_AN, rest = rest[: len(AN)], rest[len(AN):]
    inv_BM, inv_BN = rest[: len(BM)], rest[len(BM):]
    if verbose:
        print(f"There are {pairs.shape[0]} unique pole-pole combinations.")
    # pairs are sorted by (first, second): one pole source per distinct first electrode, its
    # receivers in the order of the pairs, so mini-survey data rows equal pair indices
    src_list = []
    firsts, starts = np.unique(pairs[:, 0], return_index=True)
    bounds = list(starts) + [len(pairs)]
    for f, s0, s1 in zip(firsts, bounds[:-1], bounds[1:]):
        rx = receivers.Pole(elecs[pairs[s0:s1, 1]])
        src_list.append(sources.Pole([rx], elecs[f]))
    return [dipole_rx, dipole_tx], [inv_AM, inv_AN, inv_BM, inv_BN], Survey(src_list)


def mini_survey_data(dipoles, invs, d_mini):
    """Full-survey data from mini-survey data (simulation.py:296-304)."""
    out = np.array(d_mini[invs[0]], dtype=float)
    out[dipoles[0]] -= d_mini[invs[1]]
    out[dipoles[1]] -= d_mini[invs[2]]
    out[dipoles[0] & dipoles[1]] += d_mini[invs[3]]
    return out


def mini_survey_dataT(dipoles, invs, n_mini, v):
    """Adjoint of ``mini_survey_data`` (simulation.py:306-318)."""
    out = np.zeros(n_mini)
    np.add.at(out, invs[0], v)
    np.subtract.at(out, invs[1], v[dipoles[0]])
    np.subtract.at(out, invs[2], v[dipoles[1]])
    np.add.at(out, invs[3], v[dipoles[0] & dipoles[1]])
    return out
