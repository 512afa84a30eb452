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


# ---- vector-P2 traction inner(t, v)*ds ---------------------------------------------------
def _comp_of(V):
    cd = V.cell_dofs()
    comp = np.zeros(V.dim(), int)
    for i in range(3):
        comp[cd[:, i * 10:(i + 1) * 10].ravel()] = i
    return comp


def test_vp2_traction_vs_oracle_and_known_answer(oracle):
    from dolfin_amd import workloads
    mesh, V, mu, lmbda, f3 = workloads.vp2_elasticity_3d(3, seed=4)
    t3 = np.array([0.7, -1.1, 2.0])
    facets = dfa.exterior_facets(V)
    comp = _comp_of(V)
    for marker in (None, lambda p: p[:, 0] > 1 - 1e-12):
        L = dfa.facet_traction_form(V, t3, marker=marker)
        b = dfa.assemble(L).get_local()
        mask = None if marker is None else dfa.mark_exterior_facets(V, marker)
        _, ref = oracle.assemble_facets(-1, oracle.VP2_FACET_TRACTION, mesh.coordinates(),
                                        mesh.cells(), V.cell_dofs(), facets, ndofs=V.dim(),
                                        g=t3, facet_mask=mask)
        assert _rel(b, ref) < TOL
        area = 6.0 if marker is None else 1.0
        for i in range(3):
            assert abs(b[comp == i].sum() - area * t3[i]) < 1e-12


def test_vp2_system_with_traction_and_dirichlet_vs_oracle(oracle):
    # body force + traction on x = 1, clamped on x = 0: the facet term has no matrix part, so
    # merging it into the cell tensors before apply_bc adds it to the non-Dirichlet rows only
    from dolfin_amd import workloads
    mesh, V, mu, lmbda, f3 = workloads.vp2_elasticity_3d(3, seed=6)
    Q = V.coefficient_space()
    a = dfa.elasticity_form(V, workloads.to_function(Q, mu), workloads.to_function(Q, lmbda))
    t3 = np.array([0.0, 0.0, -3.0])
    right = lambda p: p[:, 0] > 1 - 1e-12
    L = dfa.body_force_form(V, f3) + dfa.facet_traction_form(V, t3, marker=right)
    bdofs = dfa.fem.boundary_facet_nodes(V)
    xd = np.empty((V.dim(), 3))
    cd = V.cell_dofs()
    x = mesh.coordinates()
    for i in range(3):
        for k in range(10):
            xd[cd[:, i * 10 + k]] = x[mesh.cells()[:, k]] if k < 4 else 0.5 * (
                x[mesh.cells()[:, [2, 1, 1, 0, 0, 0][k - 4]]] + x[mesh.cells()[:, [3, 3, 2, 3, 2, 1][k - 4]]])
    left = bdofs[np.abs(xd[bdofs, 0]) < 1e-12]
    bc = dfa.DirichletBC(V, 0.0, left)
    A, b = dfa.assemble_system(a, L, bc)
    rp, col, vals = A.data()
    mask = np.zeros(V.dim(), np.uint8)
    mask[left] = 1
    ra, rb = oracle.assemble_system(oracle.VP2_ELASTICITY, oracle.VP2_BODY_FORCE,
                                    mesh.coordinates(), mesh.cells(), cd, rp, col, mask,
                                    np.zeros(V.dim()), coeff_dofs=Q.cell_dofs(), a_c0=mu,
                                    a_c1=lmbda, f3=f3)
    facets = dfa.exterior_facets(V)
    fm = dfa.mark_exterior_facets(V, right)
    _, bt = oracle.assemble_facets(-1, oracle.VP2_FACET_TRACTION, mesh.coordinates(),
                                   mesh.cells(), cd, facets, ndofs=V.dim(), g=t3, facet_mask=fm)
    bt[mask == 1] = 0.0
    assert _rel(vals, ra) < TOL
    assert _rel(b.get_local(), rb + bt) < TOL
