"""Pins the FP64 CPU oracle (oracle/dc_nodal.py) with checks that do not restate its formulas.

The reference holds no golden arrays for Simulation3DNodal, so the oracle is pinned by:
  * the reference's own tolerance tests: half-space pole-dipole data within 5% of the closed
    form 2 I/(4 pi sigma)(1/r_AM - 1/r_AN) (tests/em/static/test_DC_analytic.py:10-81,
    analytics/DC.py:5-38) for Neumann and Robin; e^T M_e e / 3 = total volume
    (tests/base/test_base_pde_sim.py:126-133); top faces are exactly the z = max boundary faces
    (tests/em/static/test_DC_Boundaries.py:35-46);
  * derivative identities: Taylor O(h^2) and the adjoint identity w^T J v = v^T J^T w
    (tests/em/static/test_DC_jvecjtvecadj.py:254-274);
  * an independent brute-force loop assembly of A over edges and boundary faces;
  * mathematics the formulation fixes: symmetry, A 1 = Robin diagonal, homogeneity in sigma,
    u(c sigma) = u(sigma)/c, reciprocity, dense LU on tiny meshes, closed-form alpha.
"""

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import OracleDC, OracleExpMap, OracleIdentityMap, OracleMesh
from oracle.dc_nodal import robin_boundary_matrix


def _small_mesh():
    return OracleMesh([[(10.0, 2, -1.3), (10.0, 5), (10.0, 2, 1.3)],
                       [(10.0, 2, -1.3), (10.0, 4), (10.0, 2, 1.3)],
                       [(10.0, 2, -1.3), (10.0, 3)]], origin="CCN")


def _blocky_sigma(mesh, seed=0):
    rng = np.random.default_rng(seed)
    return np.exp(np.log(0.01) + rng.standard_normal(mesh.n_cells))


def _sources(mesh):
    A = np.array([[-20.0, 0, 0], [0.0, 0, 0]])
    M = np.array([[10.0, 0, 0], [20.0, 10, 0]])
    N = np.array([[20.0, 0, 0], [30.0, 10, 0]])
    return [
        (A, np.r_[1.0, -1.0], [dict(m=M, n=N)]),
        (np.array([[10.0, -10, 0]]), np.r_[1.0], [dict(m=np.array([[-10.0, 0, 0]]), n=None),
                                                   dict(m=M, n=N)]),
    ]


# ---------------------------------------------------------------- reference tolerance tests


@pytest.mark.parametrize("bc", ["Neumann", "Robin"])
def test_halfspace_pole_dipole_analytic(bc):
    cs, npad = 25.0, 7
    hx = [(cs, npad, -1.3), (cs, 21), (cs, npad, 1.3)]
    hz = [(cs, npad, -1.3), (cs, 20)]
    mesh = OracleMesh([hx, hx, hz], origin="CCN")
    x = mesh._cc(0)
    x = x[(x > -100) & (x < 100)]
    y = mesh._cc(1)
    y = y[(y > -100) & (y < 100)]
    M = OracleMesh._ndgrid(x - 25.0, y, np.r_[0.0])
    N = OracleMesh._ndgrid(x + 25.0, y, np.r_[0.0])
    A, B = np.r_[-200.0, 0, 0], np.r_[200.0, 0, 0]

    def pd(tx):
        rAM = np.linalg.norm(M - tx, axis=1)
        rAN = np.linalg.norm(N - tx, axis=1)
        return 2 * 1.0 / (4 * np.pi * 1e-2) * (1 / rAM - 1 / rAN)

    ana = pd(A) - pd(B)
    sim = OracleDC(mesh, [(np.vstack([A, B]), np.r_[1.0, -1.0], [dict(m=M, n=N)])],
                   mapping=OracleIdentityMap(), bc_type=bc)
    d = sim.dpred(np.full(mesh.n_cells, 1e-2))
    err = np.sqrt((((d - ana) / ana) ** 2).sum() / len(d))
    assert err < 0.05


def test_edge_inner_product_volume():
    mesh = _small_mesh()
    Me = mesh.get_edge_inner_product(np.ones(mesh.n_cells))
    assert abs(Me.diagonal().sum() / 3 - mesh.cell_volumes.sum()) < 1e-10 * mesh.cell_volumes.sum()


def test_top_faces_are_zmax_faces():
    mesh = _small_mesh()
    _, alpha = robin_boundary_matrix(mesh)
    bf = mesh.boundary_faces
    top = bf[:, 2] == mesh.nodes_z[-1]
    assert np.all(alpha[top] == 0.0)
    assert np.all(alpha[~top] > 0.0)


# ---------------------------------------------------------------- brute force / invariants


def _brute_force_A(mesh, sigma, bc):
    """Loop over edges (conductance sum_c V_c sigma_c / 4 / h^2) and boundary faces."""
    nx, ny, nz = mesh.shape_cells
    hx, hy, hz = mesh.h
    n1x, n1y = nx + 1, ny + 1
    N = mesh.n_nodes
    s = sigma.reshape((nx, ny, nz), order="F")
    A = np.zeros((N, N))
    node = lambda i, j, k: i + n1x * (j + n1y * k)  # noqa: E731

    def add_edge(a, b, w):
        A[a, a] += w
        A[b, b] += w
        A[a, b] -= w
        A[b, a] -= w

    for k in range(nz + 1):
        for j in range(ny + 1):
            for i in range(nx + 1):
                if i < nx:
                    w = sum(s[i, jj, kk] * hx[i] * hy[jj] * hz[kk] / 4
                            for jj in (j - 1, j) for kk in (k - 1, k)
                            if 0 <= jj < ny and 0 <= kk < nz) / hx[i] ** 2
                    add_edge(node(i, j, k), node(i + 1, j, k), w)
                if j < ny:
                    w = sum(s[ii, j, kk] * hx[ii] * hy[j] * hz[kk] / 4
                            for ii in (i - 1, i) for kk in (k - 1, k)
                            if 0 <= ii < nx and 0 <= kk < nz) / hy[j] ** 2
                    add_edge(node(i, j, k), node(i, j + 1, k), w)
                if k < nz:
                    w = sum(s[ii, jj, k] * hx[ii] * hy[jj] * hz[k] / 4
                            for ii in (i - 1, i) for jj in (j - 1, j)
                            if 0 <= ii < nx and 0 <= jj < ny) / hz[k] ** 2
                    add_edge(node(i, j, k), node(i, j, k + 1), w)
    if bc == "Neumann":
        A[0, :] = 0.0
        A[0, 0] = 1.0
        return A
    # Robin faces: alpha = (r_vec . n) / r^2 from (median x, median y, top z), 0 on the top
    xn, yn, zn = mesh.nodes_x, mesh.nodes_y, mesh.nodes_z
    src = np.array([np.median(xn), np.median(yn), zn[-1]])
    xc, yc, zc = mesh._cc(0), mesh._cc(1), mesh._cc(2)
    faces = []
    for k in range(nz):
        for j in range(ny):
            faces.append(((xn[0], yc[j], zc[k]), (-1, 0, 0), hy[j] * hz[k], (0, j, k),
                          [node(0, j + b, k + c) for b in (0, 1) for c in (0, 1)]))
            faces.append(((xn[-1], yc[j], zc[k]), (1, 0, 0), hy[j] * hz[k], (nx - 1, j, k),
                          [node(nx, j + b, k + c) for b in (0, 1) for c in (0, 1)]))
    for k in range(nz):
        for i in range(nx):
            faces.append(((xc[i], yn[0], zc[k]), (0, -1, 0), hx[i] * hz[k], (i, 0, k),
                          [node(i + a, 0, k + c) for a in (0, 1) for c in (0, 1)]))
            faces.append(((xc[i], yn[-1], zc[k]), (0, 1, 0), hx[i] * hz[k], (i, ny - 1, k),
                          [node(i + a, ny, k + c) for a in (0, 1) for c in (0, 1)]))
    for j in range(ny):
        for i in range(nx):
            faces.append(((xc[i], yc[j], zn[0]), (0, 0, -1), hx[i] * hy[j], (i, j, 0),
                          [node(i + a, j + b, 0) for a in (0, 1) for b in (0, 1)]))
    for f, nrm, area, cell, nodes in faces:
        r = np.array(f) - src
        alpha = np.dot(r, nrm) / np.dot(r, r)
        for n in nodes:
            A[n, n] += 0.25 * alpha * area * s[cell]
    return A


@pytest.mark.parametrize("bc", ["Robin", "Neumann"])
def test_operator_equals_brute_force_loop(bc):
    mesh = OracleMesh([[(10.0, 1, -1.3), (10.0, 3), (10.0, 1, 1.3)],
                       [(7.0, 3), (9.0, 1)], [(10.0, 1, -1.3), (5.0, 2)]], origin="CCN")
    sigma = _blocky_sigma(mesh, 3)
    sim = OracleDC(mesh, [], mapping=OracleIdentityMap(), bc_type=bc)
    A = sim.getA(sigma).toarray()
    B = _brute_force_A(mesh, sigma, bc)
    assert np.abs(A - B).max() < 1e-13 * np.abs(B).max()


def test_operator_invariants():
    mesh = _small_mesh()
    sigma = _blocky_sigma(mesh)
    sim = OracleDC(mesh, [], mapping=OracleIdentityMap(), bc_type="Robin")
    A = sim.getA(sigma)
    assert abs(A - A.T).max() < 1e-15 * abs(A).max()
    b = sim.AvgBC @ sigma
    assert np.allclose(A @ np.ones(mesh.n_nodes), b, rtol=0, atol=1e-14 * abs(A).max())
    assert np.all(b >= 0) and b.max() > 0
    assert abs(sim.getA(3.0 * sigma) - 3.0 * A).max() < 1e-14 * abs(A).max()
    ev = np.linalg.eigvalsh(A.toarray())
    assert ev.min() > 0
    simN = OracleDC(mesh, [], mapping=OracleIdentityMap(), bc_type="Neumann")
    AN = simN.getA(sigma)
    r = AN @ np.ones(mesh.n_nodes)
    assert r[0] == 1.0 and np.abs(r[1:]).max() < 1e-14 * abs(AN).max()


def test_alpha_closed_form():
    mesh = OracleMesh([[(10.0, 4)], [(10.0, 2)], [(10.0, 2)]], origin=[0.0, 0.0, -20.0])
    _, alpha = robin_boundary_matrix(mesh)
    # source point (20, 10, 0); x = 0 face of cell (0, 0, 0) at (0, 5, -15), normal -x
    bf = mesh.boundary_faces
    k = np.where(np.all(np.isclose(bf, [0.0, 5.0, -15.0]), axis=1))[0][0]
    r = np.array([0.0, 5.0, -15.0]) - np.array([20.0, 10.0, 0.0])
    assert np.isclose(alpha[k], np.dot(r, [-1, 0, 0]) / np.dot(r, r))


# ---------------------------------------------------------------- solves and derivatives


def test_dense_solve_matches_splu():
    mesh = OracleMesh([[(10.0, 4)], [(10.0, 3)], [(10.0, 3)]], origin="CCN")
    srcs = [(np.array([[-10.0, 0, 0], [10.0, 0, 0]]), np.r_[1.0, -1.0],
             [dict(m=np.array([[0.0, 0, 0]]), n=None)])]
    sim = OracleDC(mesh, srcs, mapping=OracleIdentityMap())
    sigma = _blocky_sigma(mesh, 1)
    u = sim.fields(sigma)
    ud = np.linalg.solve(sim.getA(sigma).toarray(), sim.source_term())
    assert np.allclose(u, ud, rtol=1e-12, atol=1e-14 * np.abs(ud).max())


def test_scaling_and_reciprocity():
    mesh = _small_mesh()
    sigma = _blocky_sigma(mesh, 2)
    a, b = np.array([-20.0, 0.0, 0.0]), np.array([0.0, 10.0, 0.0])
    m, n = np.array([10.0, 0.0, 0.0]), np.array([20.0, -10.0, 0.0])
    fwd = OracleDC(mesh, [(np.vstack([a, b]), np.r_[1.0, -1.0], [dict(m=m[None], n=n[None])])],
                   mapping=OracleIdentityMap())
    rev = OracleDC(mesh, [(np.vstack([m, n]), np.r_[1.0, -1.0], [dict(m=a[None], n=b[None])])],
                   mapping=OracleIdentityMap())
    d1 = fwd.dpred(sigma)
    d2 = rev.dpred(sigma)
    assert np.allclose(d1, d2, rtol=1e-10)
    fwd2 = OracleDC(mesh, fwd.sources, mapping=OracleIdentityMap())
    assert np.allclose(fwd2.dpred(5.0 * sigma), d1 / 5.0, rtol=1e-10)


def _dipole_sources(mesh):
    src = _sources(mesh)
    return [(src[0][0], src[0][1], [src[0][2][0]]), (src[1][0], src[1][1], [src[1][2][1]])]


@pytest.mark.parametrize("bc", ["Robin", "Neumann"])
def test_jvec_taylor_and_adjoint(bc):
    """Second-order Taylor decay and w^T J v = v^T J^T w.  For Neumann this holds for dipole
    receivers only (see test_neumann_pole_receiver_derivative_shift)."""
    mesh = _small_mesh()
    srcs = _sources(mesh) if bc == "Robin" else _dipole_sources(mesh)
    sim = OracleDC(mesh, srcs, mapping=OracleExpMap(), bc_type=bc)
    rng = np.random.default_rng(42)
    m0 = np.log(0.01) + 0.5 * rng.standard_normal(mesh.n_cells)
    dm = rng.standard_normal(mesh.n_cells)
    d0 = sim.dpred(m0)
    Jdm = sim.Jvec(m0, dm)
    e2 = []
    for h in (1e-1, 1e-2, 1e-3):
        d = OracleDC(mesh, srcs, mapping=OracleExpMap(), bc_type=bc).dpred(m0 + h * dm)
        e2.append(np.linalg.norm(d - d0 - h * Jdm))
    order2 = np.log10(e2[0] / e2[1]), np.log10(e2[1] / e2[2])
    assert min(order2) > 1.8, order2
    w = rng.standard_normal(sim.nD)
    v = rng.standard_normal(mesh.n_cells)
    wJv = w @ sim.Jvec(m0, v)
    vJtw = v @ sim.Jtvec(m0, w)
    assert abs(wJv - vJtw) < 1e-10 * max(abs(wJv), abs(vJtw))


def test_neumann_pole_receiver_derivative_shift():
    """Neumann: the reference replaces row 0 of A by e_0 (simulation.py:535-540) but
    getADeriv (557-571) differentiates the unmodified G^T M G.  A^-1 e_0 is the constant
    vector, so Jvec differs from the true derivative by -(dA(v) u)_0 on every node: zero for
    dipole data, -(dA(v) u)_0 for pole data."""
    mesh = _small_mesh()
    srcs = _sources(mesh)
    sim = OracleDC(mesh, srcs, mapping=OracleExpMap(), bc_type="Neumann")
    rng = np.random.default_rng(9)
    m0 = np.log(0.01) + 0.5 * rng.standard_normal(mesh.n_cells)
    v = rng.standard_normal(mesh.n_cells)
    u = sim.fields(m0)
    Jv = sim.Jvec(m0, v, u=u)
    h = 1e-5
    fd = (OracleDC(mesh, srcs, mapping=OracleExpMap(), bc_type="Neumann").dpred(m0 + h * v)
          - OracleDC(mesh, srcs, mapping=OracleExpMap(), bc_type="Neumann").dpred(m0 - h * v)) / (2 * h)
    sim.set_model(m0)
    shift = []
    for s, (_, _, rxs) in enumerate(srcs):
        c = sim.getADeriv(u[:, s], v)[0]
        for rx in rxs:
            k = rx["m"].shape[0]
            shift += [-c if rx.get("n") is None else 0.0] * k
    assert np.allclose(Jv - fd, shift, rtol=1e-5, atol=1e-8 * np.abs(fd).max())


def test_getJ_and_jtj_diag():
    mesh = _small_mesh()
    sim = OracleDC(mesh, _sources(mesh), mapping=OracleExpMap())
    rng = np.random.default_rng(7)
    m0 = np.log(0.01) + 0.3 * rng.standard_normal(mesh.n_cells)
    J = sim.getJ(m0)
    v = rng.standard_normal(mesh.n_cells)
    w = rng.standard_normal(sim.nD)
    assert np.allclose(J @ v, sim.Jvec(m0, v), rtol=1e-10, atol=0)
    assert np.allclose(J.T @ w, sim.Jtvec(m0, w), rtol=1e-10, atol=1e-14 * np.abs(J).max())
    W = sp.diags(rng.random(sim.nD) + 0.5)
    diag = sim.getJtJdiag(m0, W=W)
    dense = np.diag(J.T @ (W.toarray() ** 2) @ J)
    assert np.allclose(diag, dense, rtol=1e-10)


def test_aderiv_is_derivative_of_A():
    mesh = _small_mesh()
    sim = OracleDC(mesh, [], mapping=OracleExpMap())
    rng = np.random.default_rng(5)
    m0 = np.log(0.01) + 0.3 * rng.standard_normal(mesh.n_cells)
    sim.set_model(m0)
    u = rng.standard_normal(mesh.n_nodes)
    v = rng.standard_normal(mesh.n_cells)
    h = 1e-6
    fd = (sim.getA(np.exp(m0 + h * v)) @ u - sim.getA(np.exp(m0 - h * v)) @ u) / (2 * h)
    assert np.allclose(sim.getADeriv(u, v), fd, rtol=1e-6, atol=1e-9 * np.abs(fd).max())
    lam = rng.standard_normal(mesh.n_nodes)
    lhs = lam @ sim.getADeriv(u, v)
    rhs = v @ sim.getADeriv(u, lam, adjoint=True)
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)


def test_neumann_gauge_row():
    mesh = _small_mesh()
    srcs = _sources(mesh)
    sim = OracleDC(mesh, srcs, mapping=OracleIdentityMap(), bc_type="Neumann")
    sigma = _blocky_sigma(mesh, 4)
    u = sim.fields(sigma)
    q = sim.source_term()
    assert np.allclose(u[0], q[0])


# ---------------------------------------------------------------- interpolation and widths
# The source term (sources.py:84-86) and the receiver projections (receivers.py:414-415, 493)
# both use discretize's trilinear node interpolation.  These checks fail for swapped left/right
# weights, an axis mix-up or a reversed padding order.


def _stretched_mesh():
    return OracleMesh([[(7.0, 3, -1.4), (5.0, 4), (6.0, 2, 1.25)],
                       [(4.0, 2, -1.5), (3.0, 5), (4.0, 3, 1.3)],
                       [(8.0, 3, -1.2), (2.5, 4)]], origin="CCN")


def test_unpack_widths_hand_computed():
    """(h, n, f) tuples: h*|f|^k for k = 1..n, reversed for f < 0 (the padding grows away from
    the core on both sides)."""
    from oracle import unpack_widths

    got = unpack_widths([(10.0, 3, -1.3), (5.0, 2), (10.0, 2, 1.5), 4.0])
    want = [16.9, 13.0, 10.0 * 1.3, 5.0, 5.0, 15.0, 22.5, 4.0]
    want[0] = 10.0 * 1.3 ** 3
    want[1] = 10.0 * 1.3 ** 2
    assert np.allclose(got, want, rtol=1e-14)
    mesh = _stretched_mesh()
    assert np.allclose(mesh.h[0][:3], [7.0 * 1.4 ** 3, 7.0 * 1.4 ** 2, 7.0 * 1.4])
    assert np.allclose(mesh.h[1][-3:], [4.0 * 1.3, 4.0 * 1.69, 4.0 * 1.3 ** 3])
    # origin "C" centres the axis, "N" puts its top at 0
    assert np.isclose(mesh.nodes_x[0] + mesh.nodes_x[-1], 0.0, atol=1e-12)
    assert np.isclose(mesh.nodes_z[-1], 0.0, atol=1e-12)


def test_interpolation_reproduces_trilinear_functions():
    """At random points (off nodes and off cell mid-points) on a stretched mesh, P reproduces
    1, x, y, z and the trilinear products xy, yz, xz, xyz exactly, with 8 non-negative
    weights per row."""
    mesh = _stretched_mesh()
    rng = np.random.default_rng(11)
    lo = np.array([mesh.nodes_x[0], mesh.nodes_y[0], mesh.nodes_z[0]])
    hi = np.array([mesh.nodes_x[-1], mesh.nodes_y[-1], mesh.nodes_z[-1]])
    pts = lo + (hi - lo) * rng.random((40, 3))
    P = mesh.get_interpolation_matrix(pts)
    nodes = mesh.nodes
    x, y, z = nodes.T
    for f_nodes, f_pts in [(np.ones(len(x)), np.ones(len(pts))),
                           (x, pts[:, 0]), (y, pts[:, 1]), (z, pts[:, 2]),
                           (x * y, pts[:, 0] * pts[:, 1]), (y * z, pts[:, 1] * pts[:, 2]),
                           (x * z, pts[:, 0] * pts[:, 2]),
                           (x * y * z, pts[:, 0] * pts[:, 1] * pts[:, 2])]:
        scale = np.abs(f_nodes).max()
        assert np.allclose(P @ f_nodes, f_pts, rtol=0, atol=1e-12 * scale)
    assert P.min() >= 0.0
    assert np.all(np.diff(P.indptr) == 8)


def test_interpolation_quarter_cell_weights():
    """A point a quarter of a cell past node (i, j, k) along x, on the node's y and z lines,
    weighs node i with 0.75 and node i+1 with 0.25 and nothing else; three quarters along y and
    a tenth along z give the tensor product of the 1D weights on the right 8 nodes."""
    mesh = _stretched_mesh()
    xn, yn, zn = mesh.nodes_x, mesh.nodes_y, mesh.nodes_z
    n1x, n1y = len(xn), len(yn)
    i, j, k = 4, 6, 2
    node = lambda a, b, c: a + n1x * (b + n1y * c)  # noqa: E731
    p = np.array([[xn[i] + 0.25 * (xn[i + 1] - xn[i]), yn[j], zn[k]]])
    row = mesh.get_interpolation_matrix(p).toarray()[0]
    assert np.isclose(row[node(i, j, k)], 0.75) and np.isclose(row[node(i + 1, j, k)], 0.25)
    assert np.isclose(row.sum(), 1.0) and np.count_nonzero(row) == 2
    p = np.array([[xn[i] + 0.25 * (xn[i + 1] - xn[i]), yn[j] + 0.75 * (yn[j + 1] - yn[j]),
                   zn[k] + 0.1 * (zn[k + 1] - zn[k])]])
    row = mesh.get_interpolation_matrix(p).toarray()[0]
    want = np.zeros_like(row)
    for a, wa in ((0, 0.75), (1, 0.25)):
        for b, wb in ((0, 0.25), (1, 0.75)):
            for c, wc in ((0, 0.9), (1, 0.1)):
                want[node(i + a, j + b, k + c)] = wa * wb * wc
    assert np.allclose(row, want, rtol=0, atol=1e-14)
    # a point on a node gets weight 1 there, including the last node of an axis
    for q in [(i, j, k), (n1x - 1, n1y - 1, len(zn) - 1), (0, 0, 0)]:
        p = np.array([[xn[q[0]], yn[q[1]], zn[q[2]]]])
        row = mesh.get_interpolation_matrix(p).toarray()[0]
        assert np.isclose(row[node(*q)], 1.0) and np.isclose(row.sum(), 1.0)


# ---------------------------------------------------------------- resistivity models


def test_rho_model_derivative_leaves_out_the_robin_term():
    """With a resistivity model (rhoMap) the reference's getADeriv adds no boundary term
    (resistivity/simulation.py:566-570), so its Jvec differs from the derivative of dpred by
    exactly -P A^-1 (u * AvgBC sigma' v); with Neumann it is the derivative itself.  The adjoint
    identity holds in both."""
    mesh = _small_mesh()
    srcs = _dipole_sources(mesh)
    rng = np.random.default_rng(21)
    rho0 = np.exp(-(np.log(0.01) + 0.4 * rng.standard_normal(mesh.n_cells)))
    v = rho0 * rng.standard_normal(mesh.n_cells)
    for bc in ("Robin", "Neumann"):
        sim = OracleDC(mesh, srcs, mapping=OracleIdentityMap(), bc_type=bc, param="rho")
        u = sim.fields(rho0)
        Jv = sim.Jvec(rho0, v, u=u)
        h = 1e-6
        fd = (OracleDC(mesh, srcs, mapping=OracleIdentityMap(), bc_type=bc, param="rho")
              .dpred(rho0 + h * v)
              - OracleDC(mesh, srcs, mapping=OracleIdentityMap(), bc_type=bc, param="rho")
              .dpred(rho0 - h * v)) / (2 * h)
        sim.set_model(rho0)
        if bc == "Robin":
            dsig = sim.sigma_deriv @ v
            corr = []
            for s, plist in enumerate(sim._P):
                du = sim.solve(-(u[:, s] * (sim.AvgBC @ dsig)))
                corr += [P @ du for P in plist]
            corr = np.concatenate(corr)
            assert np.linalg.norm(corr) > 1e-3 * np.linalg.norm(fd)  # the term is not negligible
            assert np.allclose(Jv + corr, fd, rtol=1e-5, atol=1e-7 * np.abs(fd).max())
        else:
            assert np.allclose(Jv, fd, rtol=1e-5, atol=1e-7 * np.abs(fd).max())
        w = rng.standard_normal(sim.nD)
        a, b = w @ Jv, v @ sim.Jtvec(rho0, w, u=u)
        assert abs(a - b) < 1e-10 * max(abs(a), abs(b))
