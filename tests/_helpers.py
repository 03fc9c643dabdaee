"""Shared test helpers: build the oracle problem that matches a simpeg_amd simulation."""

import numpy as np

from oracle import OracleDC, OracleExpMap, OracleIdentityMap, OracleMesh
from simpeg_amd.resistivity import receivers as Rx


def oracle_sources(survey):
    out = []
    for src in survey.source_list:
        locs, cur = src.electrodes()
        rxs = []
        for rx in src.receiver_list:
            if isinstance(rx, Rx.Dipole):
                d = dict(m=rx.locations[0], n=rx.locations[1])
            else:
                d = dict(m=rx.locations, n=None)
            if rx.data_type == "apparent_resistivity":
                d["factor"] = 1.0 / np.asarray(rx.geometric_factor[src])
            rxs.append(d)
        out.append((locs, cur, rxs))
    return out


def oracle_for(mesh, survey, bc_type="Robin", mapping="exp", surface_faces=None, param="sigma"):
    om = OracleMesh(mesh.h, origin=mesh.origin)
    mp = OracleExpMap() if mapping == "exp" else OracleIdentityMap()
    return OracleDC(om, oracle_sources(survey), mapping=mp, bc_type=bc_type,
                    surface_faces=surface_faces, param=param)


def rel(a, b):
    a = np.asarray(a, dtype=float)
    b = np.asarray(b, dtype=float)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def tiny_problem(seed=0, core=(10, 8, 6), npad=3, nsrc=4):
    """A small padded mesh with a blocky log-conductivity and one dipole-dipole line."""
    from simpeg_amd import workloads as W

    mesh = W.padded_mesh(core, npad=npad)
    rng = np.random.default_rng(seed)
    nx, ny, nz = mesh.shape_cells
    m = np.full((nx, ny, nz), np.log(0.01))
    m[npad + 1:npad + 4, npad + 2:npad + 5, npad:npad + 3] = np.log(0.1)
    m[npad + 5:npad + 8, npad + 1:npad + 4, npad + 1:npad + 4] = np.log(0.002)
    m = m.ravel(order="F") + 0.1 * rng.standard_normal(mesh.n_cells)
    half = core[0] * 10 / 2 - 10
    src = W.line_sources("dipole-dipole", [-half, half, 0.0, 0.0], 10.0, 3)[:nsrc]
    from simpeg_amd.resistivity.survey import Survey
    return mesh, Survey(src), m


def mixed_survey():
    """A dipole-dipole line plus pole-dipole, pole-pole and dipole-pole sources sharing
    electrodes (the mixture of tests/em/static/test_DC_miniaturize.py:230-270)."""
    from simpeg_amd import workloads as W
    from simpeg_amd.resistivity import receivers as Rx
    from simpeg_amd.resistivity import sources as Src
    from simpeg_amd.resistivity.survey import Survey

    srcs = W.line_sources("dipole-dipole", [-40.0, 40.0, 0.0, 0.0], 10.0, 3)
    e = Survey(srcs).unique_electrode_locations
    rx_p = Rx.Pole(e[[2]])
    rx_d = Rx.Dipole(e[[2]], e[[3]])
    srcs = srcs + [Src.Pole([rx_d], e[0]), Src.Pole([rx_p], e[0]), Src.Dipole([rx_p], e[0], e[1])]
    return Survey(srcs)
