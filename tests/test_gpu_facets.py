"""Exterior-facet integrals on the GPU against the oracle (tolerance: relative Frobenius / L2
1e-12) and the reference's boundary-measure known answers."""
import numpy as np
import pytest
import torch

import dolfin_amd as dfa

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _mesh(kind, n, rng, jitter=0.15):
    if kind == "interval":
        m = dfa.UnitIntervalMesh(n)
    elif kind == "square":
        m = dfa.UnitSquareMesh(n, n)
    else:
        m = dfa.UnitCubeMesh(n, n, n)
    if jitter:
        dfa.jitter_interior(m, jitter / n, rng)
    return m


def test_exterior_facets_match_oracle(oracle):
    mesh = dfa.UnitCubeMesh(3, 4, 2)
    V = dfa.FunctionSpace(mesh, "P", 1)
    fc, fl = dfa.exterior_facets(V)
    oc, ol = oracle.exterior_facets(mesh.cells())
    assert np.array_equal(fc, oc) and np.array_equal(fl, ol)


@pytest.mark.parametrize("kind,n,measure", [("square", 4, 4.0), ("cube", 2, 6.0)])
def test_boundary_measure_known_answer(kind, n, measure):
    # python/test/unit/fem/test_assembler.py:140-153
    mesh = _mesh(kind, n, None, jitter=0.0)
    V = dfa.FunctionSpace(mesh, "P", 1)
    one = dfa.Function(V, np.ones(V.dim()))
    b = dfa.assemble(dfa.facet_load_form(V, one))
    assert abs(float(b.values.sum()) - measure) < 1e-12


def test_marked_boundary_known_answer():
    # python/test/unit/fem/test_assembler.py:296-322
    mesh = dfa.UnitSquareMesh(8, 8)
    V = dfa.FunctionSpace(mesh, "P", 1)
    eps = 3e-16
    inside = lambda x: ((x[:, 0] < 0.5 + eps) & (x[:, 1] < eps)) | ((x[:, 1] < 0.5 + eps) & (x[:, 0] < eps))
    one = dfa.Function(V, np.ones(V.dim()))
    b = dfa.assemble(dfa.facet_load_form(V, one, marker=inside))
    assert abs(float(b.values.sum()) - 1.0) < 1e-12


@pytest.mark.parametrize("kind,n", [("interval", 30), ("square", 12), ("cube", 6)])
def test_facet_forms_vs_oracle(oracle, kind, n):
    rng = np.random.default_rng(11)
    mesh = _mesh(kind, n, rng)
    perm = rng.permutation(mesh.num_vertices())
    V = dfa.FunctionSpace(mesh, "P", 1, node_permutation=perm)
    cd = V.cell_dofs()
    g = rng.uniform(-1, 1, V.dim())
    kappa = rng.uniform(0.5, 2, V.dim())
    G = dfa.Function(V, g)
    K = dfa.Function(V, kappa)
    inside = lambda x: x[:, 0] > 0.25
    fac = oracle.exterior_facets(mesh.cells())
    mask = dfa.mark_exterior_facets(V, inside)
    rp, col = oracle.sparsity(cd, V.dim())
    # matrix: poisson + Robin term on the marked facets (one Assembler call on the sum)
    A = dfa.assemble(dfa.poisson_form(V, K) + dfa.facet_mass_form(V, 2.5, marker=inside))
    Kref = oracle.assemble_matrix(oracle.P1_POISSON, mesh.coordinates(), mesh.cells(), cd, rp, col,
                                  coeff_dofs=cd, c0=kappa)
    Mref, _ = oracle.assemble_facets(oracle.P1_FACET_MASS, -1, mesh.coordinates(), mesh.cells(), cd,
                                     fac, rp, col, coeff_dofs=cd, alpha=2.5, facet_mask=mask)
    grp, gcol = A.structure()
    assert np.array_equal(grp.cpu().numpy(), rp) and np.array_equal(gcol.cpu().numpy(), col)
    assert _rel(A.values.cpu().numpy(), Kref + Mref) <= TOL
    # vector: facet load alone, over all exterior facets
    b = dfa.assemble(dfa.facet_load_form(V, G))
    _, bref = oracle.assemble_facets(-1, oracle.P1_FACET_LOAD, mesh.coordinates(), mesh.cells(), cd,
                                     fac, ndofs=V.dim(), coeff_dofs=cd, g=g)
    assert _rel(b.values.cpu().numpy(), bref) <= TOL


@pytest.mark.parametrize("kind,n", [("square", 10), ("cube", 5)])
def test_system_with_facets_and_dirichlet_vs_oracle(oracle, kind, n):
    # SystemAssembler with Robin + Neumann terms on part of the boundary and Dirichlet values
    # on another part (facet tensors merged into the cell tensor before apply_bc)
    rng = np.random.default_rng(12)
    mesh = _mesh(kind, n, rng)
    V = dfa.FunctionSpace(mesh, "P", 1)
    cd = V.cell_dofs()
    kappa, f, g = (rng.uniform(lo, hi, V.dim()) for lo, hi in ((0.5, 2), (-1, 1), (-1, 1)))
    x = mesh.coordinates()
    bvert = oracle.boundary_vertices(x)
    dof_of_v = np.empty(mesh.num_vertices(), np.int64)
    dof_of_v[mesh.cells().ravel()] = cd.ravel()
    bcd = dof_of_v[bvert][x[bvert, 0] < 0.5]
    gval = rng.uniform(-1, 1, bcd.shape[0])
    bc = dfa.DirichletBC(V, gval, dofs=bcd)
    inside = lambda y: y[:, 0] > 0.3
    a = dfa.poisson_form(V, dfa.Function(V, kappa)) + dfa.facet_mass_form(V, 1.7, marker=inside)
    L = dfa.load_form(V, dfa.Function(V, f)) + dfa.facet_load_form(V, dfa.Function(V, g), marker=inside)
    A, b = dfa.assemble_system(a, L, bc)
    bcm = np.zeros(V.dim(), np.uint8)
    bcm[bcd] = 1
    bcv = np.zeros(V.dim())
    bcv[bcd] = gval
    rp, col = oracle.sparsity(cd, V.dim())
    fac = oracle.exterior_facets(mesh.cells())
    mask = dfa.mark_exterior_facets(V, inside)
    Aref, bref = oracle.assemble_system_facets(oracle.P1_POISSON, oracle.P1_LOAD,
                                               oracle.P1_FACET_MASS, oracle.P1_FACET_LOAD, x,
                                               mesh.cells(), cd, fac, rp, col, bcm, bcv,
                                               coeff_dofs=cd, a_c0=kappa, L_c0=f, alpha=1.7, g=g,
                                               facet_mask=mask)
    assert _rel(A.values.cpu().numpy(), Aref) <= TOL
    assert _rel(b.values.cpu().numpy(), bref) <= TOL
    # the linear form of facet terms only
    _, b2 = dfa.assemble_system(a, dfa.facet_load_form(V, dfa.Function(V, g), marker=inside), bc)
    _, b2ref = oracle.assemble_system_facets(oracle.P1_POISSON, oracle.P1_LOAD, oracle.P1_FACET_MASS,
                                             oracle.P1_FACET_LOAD, x, mesh.cells(), cd, fac, rp, col,
                                             bcm, bcv, coeff_dofs=cd, a_c0=kappa,
                                             L_c0=np.zeros(V.dim()), alpha=1.7, g=g, facet_mask=mask)
    assert _rel(b2.values.cpu().numpy(), b2ref) <= TOL
