"""Pins of the CPU oracle: the reference's own known-answer constants, closed forms and
mathematics that a plausible mistake (dropped term, wrong sign/index, transposed operand)
would break. CPU only."""
import numpy as np
import pytest


def test_system_assembler_p1_bc_unitcube4(oracle):
    # python/test/unit/fem/test_system_assembler.py:73-91
    O = oracle
    x, cells = O.box_mesh(4, 4, 4)
    rp, col = O.sparsity(cells, x.shape[0])
    bc = O.boundary_vertices(x).astype(np.uint8)
    ones = np.ones(x.shape[0])
    A, b = O.assemble_system(O.P1_POISSON, O.P1_LOAD, x, cells, cells, rp, col, bc, ones,
                             a_c0=ones, L_c0=np.full(x.shape[0], 10.0))
    assert round(O.frobenius(A) - 96.847818767384, 10) == 0
    assert round(np.linalg.norm(b) - 96.564760289080, 10) == 0
    # assemble(A) only / assemble(b) only give the same tensors (test_system_assembler.py:99-118)
    A2, _ = O.assemble_system(O.P1_POISSON, O.P1_LOAD, x, cells, cells, rp, col, bc, ones,
                              a_c0=ones, L_c0=np.full(x.shape[0], 10.0), want_b=False)
    _, b2 = O.assemble_system(O.P1_POISSON, O.P1_LOAD, x, cells, cells, rp, col, bc, ones,
                              a_c0=ones, L_c0=np.full(x.shape[0], 10.0), want_A=False)
    np.testing.assert_array_equal(A, A2)
    np.testing.assert_array_equal(b, b2)


def test_threaded_system_loop_matches_serial(oracle):
    # the multi-core timing variant: same values up to summation order, same known answers
    O = oracle
    x, cells = O.box_mesh(6, 5, 4)
    rp, col = O.sparsity(cells, x.shape[0])
    bc = O.boundary_vertices(x).astype(np.uint8)
    rng = np.random.default_rng(4)
    k, f, g = rng.uniform(0.5, 2, x.shape[0]), rng.uniform(-1, 1, x.shape[0]), rng.uniform(-1, 1, x.shape[0])
    A, b = O.assemble_system(O.P1_POISSON, O.P1_LOAD, x, cells, cells, rp, col, bc, g,
                             a_c0=k, L_c0=f)
    for t in (1, 3, 8):
        At, bt = O.assemble_system(O.P1_POISSON, O.P1_LOAD, x, cells, cells, rp, col, bc, g,
                                   a_c0=k, L_c0=f, threads=t)
        assert np.linalg.norm(At - A) <= 1e-14 * np.linalg.norm(A)
        assert np.linalg.norm(bt - b) <= 1e-14 * np.linalg.norm(b)


def test_interval_p1(oracle):
    # python/test/unit/fem/test_assembler.py:39-55
    O = oracle
    x, cells = O.interval_mesh(48)
    rp, col = O.sparsity(cells, x.shape[0])
    A = O.assemble_matrix(O.P1_POISSON, x, cells, cells, rp, col, c0=np.ones(x.shape[0]))
    b = O.assemble_vector(O.P1_LOAD, x, cells, cells, x.shape[0], c0=np.full(x.shape[0], 10.0))
    assert round(O.frobenius(A) - 811.75365721381274397572, 10) == 0
    assert round(np.linalg.norm(b) - 1.43583841167606474087, 10) == 0


def test_vector_dg1_epsilon(oracle):
    # python/test/unit/fem/test_assembler.py:58-77 (vector DG1, eps(v):eps(u) = mu 1/2, lmbda 0)
    O = oracle
    x, cells = O.box_mesh(4, 4, 4)
    nc = cells.shape[0]
    cd = np.arange(nc * 12, dtype=np.int32).reshape(nc, 12)
    rp, col = O.sparsity(cd, nc * 12)
    A = O.assemble_matrix(O.VP1_ELASTICITY, x, cells, cd, rp, col, c0=np.full(x.shape[0], 0.5),
                          c1=np.zeros(x.shape[0]))
    b = O.assemble_vector(O.VP1_BODY_FORCE, x, cells, cd, nc * 12, f3=np.array([10., 20., 30.]))
    assert round(O.frobenius(A) - 4.3969686527582512, 10) == 0
    assert round(np.linalg.norm(b) - 0.95470326978246278, 10) == 0


def test_reference_tetrahedron(oracle):
    # python/test/unit/fem/test_assembler.py:351-390 with fixture tetrahedron.xml
    xc = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1.]])
    A0 = np.array([[.5, -1 / 6, -1 / 6, -1 / 6], [-1 / 6, 1 / 6, 0, 0], [-1 / 6, 0, 1 / 6, 0],
                   [-1 / 6, 0, 0, 1 / 6]])
    M0 = np.full((4, 4), 1 / 120) + np.eye(4) / 120
    np.testing.assert_allclose(oracle.p1_stiffness(xc, np.ones(4)), A0, atol=1e-15)
    np.testing.assert_allclose(oracle.p1_mass(xc), M0, atol=1e-16)


def test_interval_local_tensors(oracle):
    # python/test/unit/fem/test_local_assembler.py:28-58 (UnitIntervalMesh(20), cell 0)
    xc = np.array([[0.0], [0.05]])
    np.testing.assert_allclose(oracle.p1_mass(xc), [[1 / 60, 1 / 120], [1 / 120, 1 / 60]],
                               rtol=1e-12)
    np.testing.assert_allclose(oracle.p1_load(xc, np.ones(2)), [0.025, 0.025], rtol=1e-13)


def test_mesh_sizes_and_orientation(oracle):
    x, cells = oracle.box_mesh(3, 3, 3)
    assert x.shape == (64, 3) and cells.shape == (162, 4)
    assert np.all(np.diff(cells, axis=1) > 0)  # Mesh::order(): ascending vertices
    e = x[cells[:, 1:]] - x[cells[:, :1]]
    det = np.linalg.det(e)
    # Kuhn tets after ordering: half of them have negative orientation
    assert np.sum(det < 0) == cells.shape[0] // 2
    np.testing.assert_allclose(np.sum(np.abs(det)) / 6, 1.0, rtol=1e-13)
    x2, c2 = oracle.rectangle_mesh_right(64, 64)
    assert x2.shape == (4225, 2) and c2.shape == (8192, 3)


def test_sparsity_bruteforce(oracle):
    x, cells = oracle.box_mesh(2, 2, 1)
    n = x.shape[0]
    rp, col = oracle.sparsity(cells, n)
    dense = np.zeros((n, n), bool)
    for c in cells:
        dense[np.ix_(c, c)] = True
    for r in range(n):
        np.testing.assert_array_equal(col[rp[r]:rp[r + 1]], np.nonzero(dense[r])[0])
    # keep_diagonal adds (r, r) for dofs outside every cell
    rp2, col2 = oracle.sparsity(cells, n + 2, keep_diagonal=True)
    assert col2.shape[0] == col.shape[0] + 2 and col2[-1] == n + 1


@pytest.mark.parametrize("n", [1, 2, 3])
def test_vp2_sizes(oracle, n):
    x, cells = oracle.box_mesh(n, n, n)
    ce, ne = oracle.tet_edges(cells)
    nv = x.shape[0]
    cd = oracle.vp2_dofmap(cells, ce, nv, ne)
    assert cd.max() + 1 == 3 * (nv + ne)
    rp, col = oracle.sparsity(cd, 3 * (nv + ne))
    assert col.shape[0] == 9 * (230 * n ** 3 + 138 * n ** 2 + 24 * n + 1)
    # edge numbering: ascending (vmin, vmax)
    loc = np.array([[2, 3], [1, 3], [1, 2], [0, 3], [0, 2], [0, 1]])
    keys = np.stack([cells[:, loc[:, 0]], cells[:, loc[:, 1]]], -1)
    order = np.zeros((ne, 2), np.int64)
    order[ce.ravel()] = keys.reshape(-1, 2)
    assert np.all(np.diff(order[:, 0] * nv + order[:, 1]) > 0)


# ---- vector P2 element: independent quadrature -------------------------------------------
def _gauss_jacobi_tet(npts=4):
    """Collapsed-coordinate (Stroud conical product) rule on the reference tet, exact for
    degree 2*npts-1; built from Gauss-Legendre via the Duffy map."""
    g, w = np.polynomial.legendre.leggauss(npts)
    g = 0.5 * (g + 1)
    w = 0.5 * w
    pts, wts = [], []
    for i in range(npts):
        for j in range(npts):
            for k in range(npts):
                a, b, c = g[i], g[j], g[k]
                x = a * (1 - b) * (1 - c)
                y = b * (1 - c)
                z = c
                jac = (1 - b) * (1 - c) ** 2
                pts.append((x, y, z))
                wts.append(w[i] * w[j] * w[k] * jac)
    return np.array(pts), np.array(wts)


def _p2_grad(lam, G):
    """Gradients of the 10 P2 basis functions at barycentric point lam (vertices, then edges
    (2,3),(1,3),(1,2),(0,3),(0,2),(0,1)); G[k] = grad lambda_k."""
    edges = [(2, 3), (1, 3), (1, 2), (0, 3), (0, 2), (0, 1)]
    out = np.zeros((10, 3))
    for k in range(4):
        out[k] = (4 * lam[k] - 1) * G[k]
    for e, (i, j) in enumerate(edges):
        out[4 + e] = 4 * (lam[i] * G[j] + lam[j] * G[i])
    return out


def _vp2_quadrature(xc, mu, la):
    J = (xc[1:] - xc[0]).T
    Ginv = np.linalg.inv(J)
    G = np.vstack([-Ginv.sum(axis=0), Ginv])  # rows: grad lambda_k
    vol = abs(np.linalg.det(J))
    pts, wts = _gauss_jacobi_tet(4)
    K = np.zeros((30, 30))
    for (xr, yr, zr), w in zip(pts, wts):
        lam = np.array([1 - xr - yr - zr, xr, yr, zr])
        dphi = _p2_grad(lam, G)
        m = lam @ mu
        l = lam @ la
        for i in range(3):
            for a in range(10):
                for j in range(3):
                    for b in range(10):
                        v = m * ((i == j) * dphi[a] @ dphi[b] + dphi[a, j] * dphi[b, i]) \
                            + l * dphi[a, i] * dphi[b, j]
                        K[i * 10 + a, j * 10 + b] += w * vol * v
    return K


def test_vp2_element_matches_quadrature(oracle):
    rng = np.random.default_rng(7)
    xc = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1.]]) + rng.uniform(-.2, .2, (4, 3))
    mu = rng.uniform(0.5, 1.5, 4)
    la = rng.uniform(1, 10, 4)
    K = oracle.vp2_elasticity(xc, mu, la)
    Kq = _vp2_quadrature(xc, mu, la)
    np.testing.assert_allclose(K, Kq, rtol=1e-12, atol=1e-12 * np.abs(Kq).max())


def test_vp2_rigid_modes_and_load(oracle):
    rng = np.random.default_rng(3)
    xc = rng.uniform(0, 1, (4, 3))
    K = oracle.vp2_elasticity(xc, rng.uniform(.5, 1.5, 4), rng.uniform(1, 10, 4))
    np.testing.assert_allclose(K, K.T, atol=1e-12 * np.abs(K).max())
    edges = [(2, 3), (1, 3), (1, 2), (0, 3), (0, 2), (0, 1)]
    nodes = np.vstack([xc] + [0.5 * (xc[i] + xc[j]) for i, j in edges])
    modes = []
    for c in range(3):
        r = np.zeros(30)
        r[c * 10:(c + 1) * 10] = 1
        modes.append(r)
    for axis in np.eye(3):
        disp = np.cross(axis, nodes)  # infinitesimal rotation, interpolated exactly by P2
        modes.append(disp.T.ravel())
    for r in modes:
        assert np.abs(K @ r).max() < 1e-11 * np.abs(K).max()
    vol = abs(np.linalg.det(xc[1:] - xc[0])) / 6
    b = oracle.vp2_body_force(xc, np.array([0, 0, -1.]))
    np.testing.assert_allclose(b[20:24], vol / 20)
    np.testing.assert_allclose(b[24:30], -vol / 5)
    assert np.all(b[:20] == 0)
    np.testing.assert_allclose(b.sum(), -vol, rtol=1e-14)


# ---- adjoint closed forms vs central finite differences of the assembled energy ------------
def _energy_p1(O, x, cells, kappa, f, u, rp, col):
    A = O.assemble_matrix(O.P1_POISSON, x, cells, cells, rp, col, c0=kappa)
    b = O.assemble_vector(O.P1_LOAD, x, cells, cells, x.shape[0], c0=f)
    Ad = np.zeros((x.shape[0],) * 2)
    rows = np.repeat(np.arange(x.shape[0]), np.diff(rp))
    Ad[rows, col] = A
    return u @ Ad @ u - b @ u


@pytest.mark.parametrize("dim", [2, 3])
def test_adjoint_p1_fd(oracle, dim):
    O = oracle
    rng = np.random.default_rng(11)
    if dim == 2:
        x, cells = O.rectangle_mesh_right(4, 3)
    else:
        x, cells = O.box_mesh(2, 2, 2)
    x = x + rng.uniform(-0.05, 0.05, x.shape) * (1 - O.boundary_vertices(x))[:, None]
    n = x.shape[0]
    kappa, f, u = rng.uniform(.5, 2, n), rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    rp, col = O.sparsity(cells, n)
    J, dk, dx = O.adjoint_p1(x, cells, cells, kappa, f, u)
    assert abs(J - _energy_p1(O, x, cells, kappa, f, u, rp, col)) < 1e-12 * max(1, abs(J))
    h = 1e-6
    for _ in range(4):
        dkap = rng.standard_normal(n)
        fd = (_energy_p1(O, x, cells, kappa + h * dkap, f, u, rp, col)
              - _energy_p1(O, x, cells, kappa - h * dkap, f, u, rp, col)) / (2 * h)
        assert abs(fd - dk @ dkap) < 1e-6 * max(1, abs(fd))
        dX = rng.standard_normal(x.shape)
        fd = (_energy_p1(O, x + h * dX, cells, kappa, f, u, rp, col)
              - _energy_p1(O, x - h * dX, cells, kappa, f, u, rp, col)) / (2 * h)
        assert abs(fd - np.sum(dx * dX)) < 1e-6 * max(1, abs(fd))


def _energy_vp2(O, x, cells, cd, kd, mu, la, f3, u, rp, col):
    A = O.assemble_matrix(O.VP2_ELASTICITY, x, cells, cd, rp, col, coeff_dofs=kd, c0=mu, c1=la)
    b = O.assemble_vector(O.VP2_BODY_FORCE, x, cells, cd, u.shape[0], f3=f3)
    rows = np.repeat(np.arange(u.shape[0]), np.diff(rp))
    return np.sum(A * u[rows] * u[col]) - b @ u


def test_adjoint_vp2_fd(oracle):
    O = oracle
    rng = np.random.default_rng(5)
    x, cells = O.box_mesh(1, 2, 1)
    x = x + rng.uniform(-0.05, 0.05, x.shape)
    ce, ne = O.tet_edges(cells)
    nv = x.shape[0]
    cd = O.vp2_dofmap(cells, ce, nv, ne)
    nd = 3 * (nv + ne)
    rp, col = O.sparsity(cd, nd)
    mu, la, u = rng.uniform(.5, 1.5, nv), rng.uniform(1, 10, nv), rng.uniform(-1, 1, nd)
    f3 = np.array([0.3, -0.2, -1.0])
    J, dmu, dla, dx = O.adjoint_vp2(x, cells, cd, cells, mu, la, f3, u)
    E = lambda xx, m, l: _energy_vp2(O, xx, cells, cd, cells, m, l, f3, u, rp, col)
    assert abs(J - E(x, mu, la)) < 1e-11 * max(1, abs(J))
    h = 1e-6
    for _ in range(3):
        d = rng.standard_normal(nv)
        fd = (E(x, mu + h * d, la) - E(x, mu - h * d, la)) / (2 * h)
        assert abs(fd - dmu @ d) < 1e-6 * max(1, abs(fd))
        fd = (E(x, mu, la + h * d) - E(x, mu, la - h * d)) / (2 * h)
        assert abs(fd - dla @ d) < 1e-6 * max(1, abs(fd))
        dX = rng.standard_normal(x.shape)
        fd = (E(x + h * dX, mu, la) - E(x - h * dX, mu, la)) / (2 * h)
        assert abs(fd - np.sum(dx * dX)) < 1e-6 * max(1, abs(fd))


def test_shape_derivative_identity_p1(oracle):
    # test_coordinate_derivative.py:47-49: dJ[V] for J = int |grad u|^2 equals
    # int -2 (grad V grad u).grad u + |grad u|^2 div V; for P1 V the integrand is cellwise
    # constant, so the assembled form is a closed-form sum that must match the adjoint.
    O = oracle
    rng = np.random.default_rng(2)
    x, cells = O.rectangle_mesh_right(3, 3)
    n = x.shape[0]
    u = rng.uniform(-1, 1, n)
    V = rng.standard_normal(x.shape)
    J, dk, dx = O.adjoint_p1(x, cells, cells, np.ones(n), np.zeros(n), u)
    total = 0.0
    for c in cells:
        J_ = (x[c[1:]] - x[c[0]]).T
        Gi = np.linalg.inv(J_)
        G = np.vstack([-Gi.sum(0), Gi])
        vol = abs(np.linalg.det(J_)) / 2
        gu = u[c] @ G
        gV = V[c].T @ G  # grad V, gV[i, j] = d V_i / d x_j
        total += vol * (-2 * gu @ gV @ gu + gu @ gu * np.trace(gV))
    assert abs(total - np.sum(dx * V)) < 1e-12 * max(1, abs(total))


def _dense(rp, col, vals):
    n = rp.shape[0] - 1
    A = np.zeros((n, n))
    A[np.repeat(np.arange(n), np.diff(rp)), col] = vals
    return A


def test_x0_newton_increment(oracle):
    # SystemAssembler::assemble(A, b, x0) (SystemAssembler.cpp:252-275). The boundary values
    # become x0 - g, so the increment du of A du = b makes x0 - du meet the condition g (a
    # sign error would give 2 x0 - g); x0 leaves A unchanged; x0 = g gives the system of the
    # homogeneous condition.
    O = oracle
    x, cells = O.box_mesh(3, 3, 3)
    n = x.shape[0]
    rp, col = O.sparsity(cells, n)
    bc = O.boundary_vertices(x).astype(np.uint8)
    rng = np.random.default_rng(11)
    k, f = rng.uniform(0.5, 2, n), rng.uniform(-1, 1, n)
    g, x0 = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    A, b = O.assemble_system(O.P1_POISSON, O.P1_LOAD, x, cells, cells, rp, col, bc, g,
                             a_c0=k, L_c0=f)
    Ax, bx = O.assemble_system(O.P1_POISSON, O.P1_LOAD, x, cells, cells, rp, col, bc, g,
                               a_c0=k, L_c0=f, x0=x0)
    np.testing.assert_array_equal(A, Ax)
    du = np.linalg.solve(_dense(rp, col, Ax), bx)
    m = bc != 0
    np.testing.assert_allclose((x0 - du)[m], g[m], rtol=0, atol=1e-12)
    _, bg = O.assemble_system(O.P1_POISSON, O.P1_LOAD, x, cells, cells, rp, col, bc, g,
                              a_c0=k, L_c0=f, x0=np.where(m, g, 7.0))
    _, b0 = O.assemble_system(O.P1_POISSON, O.P1_LOAD, x, cells, cells, rp, col, bc,
                              np.zeros(n), a_c0=k, L_c0=f)
    np.testing.assert_array_equal(bg, b0)
    assert np.all(bg[m] == 0.0)
