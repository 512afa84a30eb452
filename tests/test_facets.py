"""Exterior-facet integrals (Neumann / Robin terms): the oracle against the reference's
known answers and exact facet integrals, the library's facet extraction against the oracle,
and the decomposition the GPU pass uses for Dirichlet rows (CPU only)."""
import numpy as np
import pytest

import dolfin_amd as dfa
from conftest import p

pytestmark = []


def _mesh(kind, n):
    if kind == "interval":
        return dfa.UnitIntervalMesh(n)
    if kind == "square":
        return dfa.UnitSquareMesh(n, n)
    return dfa.UnitCubeMesh(n, n, n)


def _facet_mask(mesh, facets, inside):
    """SubDomain::mark rule (dolfin/mesh/SubDomain.cpp:241-268): all vertices and the midpoint."""
    fc, fl = facets
    cells, x = mesh.cells(), mesh.coordinates()
    nvc = cells.shape[1]
    cols = np.array([[a for a in range(nvc) if a != f] for f in range(nvc)])
    pts = x[cells[fc[:, None], cols[fl]]]
    ok = np.ones(fc.shape[0], bool)
    for j in range(pts.shape[1]):
        ok &= inside(pts[:, j, :])
    ok &= inside(pts.mean(axis=1))
    return ok.astype(np.uint8)


@pytest.mark.parametrize("kind,n,measure", [("square", 4, 4.0), ("cube", 2, 6.0), ("interval", 5, 2.0)])
def test_boundary_measure_known_answer(oracle, kind, n, measure):
    # python/test/unit/fem/test_assembler.py:140-153: assemble(1*ds(mesh)) = 4.0 (square),
    # 6.0 (cube); here as sum of b for g*v*ds with g = 1
    mesh = _mesh(kind, n)
    V = dfa.FunctionSpace(mesh, "P", 1)
    fac = oracle.exterior_facets(mesh.cells())
    _, b = oracle.assemble_facets(-1, oracle.P1_FACET_LOAD, mesh.coordinates(), mesh.cells(),
                                  V.cell_dofs(), fac, ndofs=V.dim(), coeff_dofs=V.cell_dofs(),
                                  g=np.ones(V.dim()))
    assert abs(b.sum() - measure) < 1e-12
    expected = {"interval": 2, "square": 4 * n, "cube": 12 * n * n}[kind]
    assert fac[0].shape[0] == expected


def test_marked_boundary_known_answer(oracle):
    # python/test/unit/fem/test_assembler.py:296-322: c*ds(0) over the lower-left boundary
    # (x < 0.5, y = 0) or (y < 0.5, x = 0) of UnitSquareMesh(8, 8) is 1.0
    mesh = dfa.UnitSquareMesh(8, 8)
    V = dfa.FunctionSpace(mesh, "P", 1)
    fac = oracle.exterior_facets(mesh.cells())
    eps = 3e-16
    inside = lambda x: ((x[:, 0] < 0.5 + eps) & (x[:, 1] < eps)) | ((x[:, 1] < 0.5 + eps) & (x[:, 0] < eps))
    mask = _facet_mask(mesh, fac, inside)
    _, b = oracle.assemble_facets(-1, oracle.P1_FACET_LOAD, mesh.coordinates(), mesh.cells(),
                                  V.cell_dofs(), fac, ndofs=V.dim(), coeff_dofs=V.cell_dofs(),
                                  g=np.ones(V.dim()), facet_mask=mask)
    assert abs(b.sum() - 1.0) < 1e-12
    assert int(mask.sum()) == 8


def test_facet_tensors_exact(oracle):
    # int_F lambda_a lambda_b = |F| (1 + delta_ab) / ((k + 1)(k + 2)) on a k-simplex
    # (a! b! k! / (a + b + k)! |F| k!-normalised monomial rule): triangle facet |F| = 1/2
    xc = np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]])
    M = oracle.p1_facet_mass(xc, 3, alpha=2.0)
    ref = np.zeros((4, 4))
    ref[:3, :3] = 2.0 * 0.5 * (np.ones((3, 3)) + np.eye(3)) / 12.0
    assert np.allclose(M, ref, rtol=0, atol=1e-15)
    # facet opposite vertex 0: the slanted face, area sqrt(3)/2
    g = np.array([0.0, 1.0, 2.0, 3.0])
    b = oracle.p1_facet_load(xc, 0, g)
    area = np.sqrt(3.0) / 2.0
    Mf = area * (np.ones((3, 3)) + np.eye(3)) / 12.0
    assert np.allclose(b, np.concatenate([[0.0], Mf @ g[1:]]), rtol=1e-14, atol=0)
    # edge facet in 2D: length sqrt(2), int lambda_a lambda_b = L (1 + delta) / 6
    xt = np.array([[0.0, 0], [1, 0], [0, 1]])
    M2 = oracle.p1_facet_mass(xt, 0)
    assert np.allclose(M2[1:, 1:], np.sqrt(2.0) * (np.ones((2, 2)) + np.eye(2)) / 6.0, rtol=1e-15)


@pytest.mark.parametrize("kind,n", [("interval", 7), ("square", 5), ("cube", 3)])
def test_library_facets_match_oracle(oracle, plan_lib, kind, n):
    mesh = _mesh(kind, n)
    cells = np.ascontiguousarray(mesh.cells(), np.int32)
    fc, fl = oracle.exterior_facets(cells)
    m = plan_lib.facets_host(mesh.gdim(), mesh.num_cells(), p(cells), None, None)
    lc, ll = np.empty(m, np.int32), np.empty(m, np.int32)
    plan_lib.facets_host(mesh.gdim(), mesh.num_cells(), p(cells), p(lc), p(ll))
    assert np.array_equal(lc, fc) and np.array_equal(ll, fl)


@pytest.mark.parametrize("kind", ["square", "cube"])
def test_facet_elimination_decomposition(oracle, kind):
    # SystemAssembler eliminates Dirichlet dofs from the cell tensor WITH its facet terms
    # (SystemAssembler.cpp:421-481). The GPU runs the cell pass and then a facet pass that
    # drops Dirichlet rows and lifts Dirichlet columns; this checks on the CPU that the two
    # agree: merged == cells-only system + facet terms without the unit diagonal.
    rng = np.random.default_rng(3)
    mesh = _mesh(kind, 4)
    dfa.jitter_interior(mesh, 0.15 / 4, rng)
    V = dfa.FunctionSpace(mesh, "P", 1)
    cd = V.cell_dofs()
    rp, col = oracle.sparsity(cd, V.dim())
    kappa = rng.uniform(0.5, 2, V.dim())
    f = rng.uniform(-1, 1, V.dim())
    g = rng.uniform(-1, 1, V.dim())
    x = mesh.coordinates()
    fac = oracle.exterior_facets(mesh.cells())
    fmask = _facet_mask(mesh, fac, lambda y: y[:, 0] > 0.3)
    bcm = np.zeros(V.dim(), np.uint8)
    bvert = oracle.boundary_vertices(x)
    dof_of_v = np.empty(mesh.num_vertices(), np.int64)
    dof_of_v[mesh.cells().ravel()] = cd.ravel()
    bcd = dof_of_v[bvert][x[bvert, 0] < 0.5]
    bcm[bcd] = 1
    bcv = np.where(bcm == 1, rng.uniform(-1, 1, V.dim()), 0.0)
    A1, b1 = oracle.assemble_system_facets(oracle.P1_POISSON, oracle.P1_LOAD, oracle.P1_FACET_MASS,
                                           oracle.P1_FACET_LOAD, x, mesh.cells(), cd, fac, rp, col,
                                           bcm, bcv, coeff_dofs=cd, a_c0=kappa, L_c0=f, alpha=1.7,
                                           g=g, facet_mask=fmask)
    A0, b0 = oracle.assemble_system(oracle.P1_POISSON, oracle.P1_LOAD, x, mesh.cells(), cd, rp,
                                    col, bcm, bcv, coeff_dofs=cd, a_c0=kappa, L_c0=f)
    Mf, _ = oracle.assemble_facets(oracle.P1_FACET_MASS, -1, x, mesh.cells(), cd, fac, rp, col,
                                   coeff_dofs=cd, alpha=1.7, facet_mask=fmask)
    _, bf = oracle.assemble_facets(-1, oracle.P1_FACET_LOAD, x, mesh.cells(), cd, fac,
                                   ndofs=V.dim(), coeff_dofs=cd, g=g, facet_mask=fmask)
    rows = np.repeat(np.arange(V.dim()), np.diff(rp))
    keep = (bcm[rows] == 0) & (bcm[col] == 0)
    A2 = A0 + np.where(keep, Mf, 0.0)
    lift = np.zeros(V.dim())
    sel = (bcm[rows] == 0) & (bcm[col] == 1)
    np.add.at(lift, rows[sel], Mf[sel] * bcv[col[sel]])
    b2 = b0 + np.where(bcm == 0, bf - lift, 0.0)
    assert np.linalg.norm(A1 - A2) <= 1e-13 * np.linalg.norm(A1)
    assert np.linalg.norm(b1 - b2) <= 1e-13 * np.linalg.norm(b1)


# ---- vector-P2 traction inner(t, v)*ds ---------------------------------------------------
_EDGES = [(2, 3), (1, 3), (1, 2), (0, 3), (0, 2), (0, 1)]


def _p2_basis(lam):
    """The ten P2 basis functions of a tetrahedron at barycentric point lam (vertex nodes
    lambda(2 lambda - 1), edge nodes 4 lambda_a lambda_b)."""
    return np.array([l * (2 * l - 1) for l in lam] + [4 * lam[a] * lam[b] for a, b in _EDGES])


def test_vp2_traction_oracle_vs_midpoint_quadrature(oracle):
    # the edge-midpoint rule on a triangle is exact for quadratics: int_F phi =
    # |F| / 3 sum over the facet's edge midpoints of phi there
    rng = np.random.default_rng(11)
    t3 = np.array([0.3, -1.2, 2.5])
    for trial in range(5):
        xc = rng.uniform(-1, 1, (4, 3))
        for lf in range(4):
            fv = [a for a in range(4) if a != lf]
            e1, e2 = xc[fv[1]] - xc[fv[0]], xc[fv[2]] - xc[fv[0]]
            area = 0.5 * np.linalg.norm(np.cross(e1, e2))
            integ = np.zeros(10)
            for a, b in [(fv[0], fv[1]), (fv[0], fv[2]), (fv[1], fv[2])]:
                lam = np.zeros(4)
                lam[a] = lam[b] = 0.5
                integ += area / 3 * _p2_basis(lam)
            ref = np.concatenate([t3[i] * integ for i in range(3)])
            np.testing.assert_allclose(oracle.vp2_facet_traction(xc, lf, t3), ref, rtol=0, atol=1e-14)


def test_vp2_traction_boundary_known_answer(oracle):
    # partition of unity on every facet: the components of b sum to t_i times the boundary
    # area (6 for the unit cube), and to t_i on the marked face x = 1
    x, cells = oracle.box_mesh(3, 3, 3)
    ce, ne = oracle.tet_edges(cells)
    cd = oracle.vp2_dofmap(cells, ce, x.shape[0], ne)
    nd = 3 * (x.shape[0] + ne)
    facets = oracle.exterior_facets(cells)
    t3 = np.array([1.0, -2.0, 0.5])
    _, b = oracle.assemble_facets(-1, oracle.VP2_FACET_TRACTION, x, cells, cd, facets, ndofs=nd, g=t3)
    comp = np.zeros(nd, int)
    for i in range(3):
        comp[cd[:, i * 10:(i + 1) * 10].ravel()] = i
    for i in range(3):
        assert abs(b[comp == i].sum() - 6.0 * t3[i]) < 1e-12
    mesh = dfa.UnitCubeMesh(3, 3, 3)
    m = _facet_mask(mesh, facets, lambda p: p[:, 0] > 1 - 1e-12)
    _, bm = oracle.assemble_facets(-1, oracle.VP2_FACET_TRACTION, x, cells, cd, facets, ndofs=nd,
                                   g=t3, facet_mask=m)
    for i in range(3):
        assert abs(bm[comp == i].sum() - t3[i]) < 1e-12
