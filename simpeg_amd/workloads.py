"""Seeded synthetic workloads for the nodal DC path (meshes, models, surveys).

* ``line_sources``: dipole-dipole / pole-dipole / dipole-pole / pole-pole sources along a surface
  line, with the station layout of ``generate_dcip_sources_line``
  (electromagnetics/static/utils/static_utils.py:1265-1413; a source whose receiver count
  would be 0 is dropped).
* ``gaussian_random_field``: white noise on a uniform grid, filtered in the Fourier domain
  to a Gaussian covariance exp(-r^2 / L^2), scaled to the requested standard deviation and
  sampled at the cell centres (nearest grid point).
* ``small_block_model`` (24x24x12 core, 20 sources), ``large_grf_model`` (96x96x64 core, 128
  sources), and the misfit / uneven-batch / large-survey variants.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .mesh import TensorMesh
from .resistivity import receivers as rx_mod
from .resistivity import sources as src_mod
from .resistivity.survey import Survey


def padded_mesh(core, cell=10.0, npad=5, factor=1.3, npad_z=None):
    """Tensor mesh with a core of `core` cells, `npad` cells growing by `factor` on both lateral
    sides and below, centred in x/y with the top at z = 0."""
    nx, ny, nz = core
    npz = npad if npad_z is None else npad_z
    hx = [(cell, npad, -factor), (cell, nx), (cell, npad, factor)]
    hy = [(cell, npad, -factor), (cell, ny), (cell, npad, factor)]
    hz = [(cell, npz, -factor), (cell, nz)]
    return TensorMesh([hx, hy, hz], origin="CCN")


def line_sources(survey_type, end_points, station_spacing, num_rx_per_src, z=0.0,
                 data_type="volt"):
    """Sources of one straight surface line (3D end points [x1, x2, y1, y2])."""
    st = survey_type.lower()
    if st not in ("pole-pole", "pole-dipole", "dipole-pole", "dipole-dipole"):
        raise ValueError("survey_type must be one of pole-pole, pole-dipole, dipole-pole, dipole-dipole")
    x1, x2, y1, y2 = [float(v) for v in end_points]
    L = np.sqrt((x2 - x1) ** 2 + (y2 - y1) ** 2)
    nstn = int(np.floor(L / station_spacing) + 1)
    dlx, dly = (x2 - x1) / L, (y2 - y1) / L
    sx = x1 + np.arange(nstn) * dlx * station_spacing
    sy = y1 + np.arange(nstn) * dly * station_spacing
    P = np.c_[sx, sy, np.full(nstn, z)]
    shift = {"pole-pole": 0, "pole-dipole": 1, "dipole-pole": 1, "dipole-dipole": 2}[st]
    out = []
    for ii in range(nstn - shift):
        D = np.sqrt((sx[ii + shift] - x2) ** 2 + (sy[ii + shift] - y2) ** 2)
        nrec = int(min(np.floor(D / station_spacing), num_rx_per_src))
        if nrec <= 0:
            continue
        if st in ("dipole-pole", "pole-pole"):
            rx = rx_mod.Pole(P[ii + shift + 1: ii + shift + nrec + 1], data_type=data_type)
        else:
            rx = rx_mod.Dipole(P[ii + shift: ii + shift + nrec], P[ii + shift + 1: ii + shift + nrec + 1],
                               data_type=data_type)
        if st in ("pole-dipole", "pole-pole"):
            out.append(src_mod.Pole([rx], P[ii]))
        else:
            out.append(src_mod.Dipole([rx], P[ii], P[ii + 1]))
    return out


def gaussian_random_field(mesh, std=1.0, corr=60.0, seed=0, spacing=10.0):
    """Zero-mean Gaussian random field at the cell centres, covariance std^2 exp(-r^2/corr^2)."""
    rng = np.random.default_rng(seed)
    lo = np.array([mesh.nodes_x[0], mesh.nodes_y[0], mesh.nodes_z[0]])
    hi = np.array([mesh.nodes_x[-1], mesh.nodes_y[-1], mesh.nodes_z[-1]])
    padw = 3.0 * corr
    n = np.ceil((hi - lo + 2 * padw) / spacing).astype(int) + 1
    noise = rng.standard_normal(tuple(n))
    k = [2 * np.pi * np.fft.fftfreq(m, d=spacing) for m in n]
    K2 = k[0][:, None, None] ** 2 + k[1][None, :, None] ** 2 + k[2][None, None, :] ** 2
    # kernel exp(-2 r^2 / L^2) convolved with itself gives covariance exp(-r^2 / L^2)
    filt = np.exp(-K2 * corr ** 2 / 8.0)
    field_ = np.real(np.fft.ifftn(np.fft.fftn(noise) * filt))
    field_ = (field_ - field_.mean()) / field_.std() * std
    cc = mesh.cell_centers
    idx = np.rint((cc - (lo - padw)) / spacing).astype(int)
    idx = np.clip(idx, 0, n - 1)
    return field_[idx[:, 0], idx[:, 1], idx[:, 2]]


@dataclass
class Workload:
    name: str
    mesh: TensorMesh
    survey: Survey
    model: np.ndarray                       # log-conductivity
    info: dict = field(default_factory=dict)


def small_block_model(seed=0):
    """24x24x12 core of 10 m cells, 5 pads x1.3 (34x34x17 cells); log sigma with a 0.1 S/m and
    a 0.001 S/m block (3-6 cells per axis) in a 0.01 S/m halfspace; 2 dipole-dipole lines at
    y = +-40 m, 20 m stations, n = 1..6, 10 sources per line."""
    mesh = padded_mesh((24, 24, 12), npad=5)
    rng = np.random.default_rng(seed)
    nx, ny, nz = mesh.shape_cells
    m = np.full((nx, ny, nz), np.log(0.01))
    core_lo = np.array([5, 5, 5])
    core_hi = np.array([5 + 24, 5 + 24, 5 + 12])
    for val in (np.log(0.1), np.log(0.001)):
        size = rng.integers(3, 7, size=3)
        start = [rng.integers(core_lo[a], core_hi[a] - size[a] + 1) for a in range(3)]
        sl = tuple(slice(start[a], start[a] + size[a]) for a in range(3))
        m[sl] = val
    src = []
    for y in (-40.0, 40.0):
        src += line_sources("dipole-dipole", [-120.0, 120.0, y, y], 20.0, 6)
    return Workload("small_block_24x24x12_20src", mesh, Survey(src), m.ravel(order="F"),
                    {"core": (24, 24, 12), "npad": 5, "seed": seed})


def grid_survey(nlines, nsrc_per_line, nrx, spacing=20.0, y_span=700.0, survey_type="dipole-dipole"):
    """`nlines` parallel lines (y evenly spread over +-y_span/2), each keeping its first
    `nsrc_per_line` sources with exactly `nrx` receivers."""
    shift = 2 if survey_type == "dipole-dipole" else 1
    nstn = nsrc_per_line + shift + nrx
    length = (nstn - 1) * spacing
    x1 = -0.5 * length
    ys = np.linspace(-0.5 * y_span, 0.5 * y_span, nlines) if nlines > 1 else np.zeros(1)
    src = []
    for y in ys:
        y = float(np.round(y / 10.0) * 10.0)
        s = line_sources(survey_type, [x1, x1 + length, y, y], spacing, nrx)
        src += s[:nsrc_per_line]
    return Survey(src)


def large_grf_model(seed=0, nlines=8, nsrc_per_line=16, nrx=16):
    """96x96x64 core of 10 m cells + 8 pads x1.3 (112x112x72 cells, 932,137 nodes);
    log sigma = log(0.01) + GRF(std 1, corr 60 m); 8 lines x 16 dipole-dipole sources x 16 Rx."""
    mesh = padded_mesh((96, 96, 64), npad=8)
    m = np.log(0.01) + gaussian_random_field(mesh, std=1.0, corr=60.0, seed=seed)
    survey = grid_survey(nlines, nsrc_per_line, nrx)
    return Workload("grf_96x96x64_128src", mesh, survey, m,
                    {"core": (96, 96, 64), "npad": 8, "seed": seed})


def medium_grf_model(seed=5):
    """32x32x20 core of 10 m cells + 6 pads x1.3 (44x44x26 cells, 54,675 nodes): the largest
    mesh of the large_grf_model construction that the FP64 direct solver factors in seconds;
    log sigma = log(0.01) + GRF(std 1, corr 60 m); 4 lines x 8 dipole-dipole sources x 8 Rx
    (32 sources, 256 data)."""
    mesh = padded_mesh((32, 32, 20), npad=6)
    m = np.log(0.01) + gaussian_random_field(mesh, std=1.0, corr=60.0, seed=seed)
    survey = grid_survey(4, 8, 8, spacing=20.0, y_span=200.0)
    return Workload("grf_32x32x20_32src", mesh, survey, m,
                    {"core": (32, 32, 20), "npad": 6, "seed": seed})


def misfit_problem(seed=0):
    """Block model + survey of ``small_block_model`` with a data-misfit setup: observed data
    from a second log-conductivity (log 0.01 + GRF(std 0.5, corr 40 m, seed+1)) plus 3 % Gaussian
    noise (seed+2); weights W = 1 / (0.03 |dobs| + 1e-5).  Returns (workload, true model).
    The observed data are computed by the caller (they need a forward solve)."""
    wl = small_block_model(seed)
    m_true = np.log(0.01) + gaussian_random_field(wl.mesh, std=0.5, corr=40.0, seed=seed + 1)
    return wl, m_true


def noisy_weights(dclean, seed=2, rel=0.03, floor=1e-5):
    rng = np.random.default_rng(seed)
    dobs = dclean + rel * np.abs(dclean) * rng.standard_normal(dclean.shape)
    return dobs, 1.0 / (rel * np.abs(dobs) + floor)


def layered_ellipsoid_model(seed=0, core=(160, 160, 96), npad=8, nlines=16, nsrc_per_line=16,
                            nrx=24):
    """160x160x96 core of 10 m cells + 8 pads x1.3 (176x176x104 cells, 3,289,545 nodes);
    log sigma: layers 0.005 / 0.02 / 0.002 S/m at depths 0-100 / 100-400 / 400+ m plus three
    seeded ellipsoids of 0.2-0.5 S/m in the core; 16 lines x 16 dipole-dipole sources x 24 Rx
    (256 sources, 6,144 data)."""
    mesh = padded_mesh(core, npad=npad)
    cc = mesh.cell_centers
    depth = -cc[:, 2]
    sig = np.where(depth < 100.0, 0.005, np.where(depth < 400.0, 0.02, 0.002))
    rng = np.random.default_rng(seed)
    half = np.array(core) * 10.0 / 2.0
    for _ in range(3):
        c = np.array([rng.uniform(-0.6, 0.6) * half[0], rng.uniform(-0.6, 0.6) * half[1],
                      -rng.uniform(80.0, 0.6 * core[2] * 10.0)])
        ax = rng.uniform(40.0, 150.0, size=3)
        inside = (((cc - c) / ax) ** 2).sum(axis=1) <= 1.0
        sig[inside] = rng.uniform(0.2, 0.5)
    spacing = 20.0
    length = (nsrc_per_line + 2 + nrx - 1) * spacing
    survey = grid_survey(nlines, nsrc_per_line, nrx, spacing=spacing,
                         y_span=min(1400.0, 2 * half[1] - 100.0))
    assert length < 2 * half[0]
    return Workload("layered_160x160x96_256src", mesh, survey, np.log(sig),
                    {"core": core, "npad": npad, "seed": seed})


def uneven_batch(seed=0, n_problems=48, sides=(24, 32, 48, 64, 80), npad=4, min_src=8,
                 max_src=96):
    """48 independent problems: cube cores with side drawn from `sides` (z extent half of x),
    4 pads x1.3, 8-96 pole-dipole sources on random surface lines, log sigma = log 0.01 +
    GRF(std 0.8, corr 40 m).  Returns a list of Workloads."""
    rng = np.random.default_rng(seed)
    out = []
    for p in range(n_problems):
        side = int(rng.choice(sides))
        mesh = padded_mesh((side, side, side // 2), npad=npad)
        m = np.log(0.01) + gaussian_random_field(mesh, std=0.8, corr=40.0, seed=seed * 1000 + p)
        nsrc = int(rng.integers(min_src, max_src + 1))
        half = side * 10.0 / 2.0
        src = []
        nl = 0
        while len(src) < nsrc and nl < 64:
            y = float(np.round(rng.uniform(-0.8, 0.8) * half / 10.0) * 10.0)
            line = line_sources("pole-dipole", [-0.9 * half, 0.9 * half, y, y], 10.0, 8)
            src += line[: nsrc - len(src)]
            nl += 1
        out.append(Workload(f"batch{p}_{side}", mesh, Survey(src), m,
                            {"side": side, "nsrc": len(src)}))
    return out
