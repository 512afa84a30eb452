"""GPU parity on unstructured simplicial meshes (Delaunay triangulations of seeded random
points plus the domain's corners, cells ordered as DOLFIN orders them). Vertex degrees vary
(up to ~30 neighbours in 3D), so CSR rows are longer than on the structured meshes and the
plan builders take their general paths. Bar as elsewhere: structure bit-exact, values 1e-12
relative (Frobenius / L2)."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.spatial
import torch

import dolfin_amd as dfa
from dolfin_amd import workloads

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _delaunay(gdim, n, seed):
    rng = np.random.default_rng(seed)
    corners = np.array(np.meshgrid(*[[0.0, 1.0]] * gdim, indexing="ij")).reshape(gdim, -1).T
    pts = np.vstack([rng.uniform(0, 1, (n, gdim)), corners])
    cells = np.sort(scipy.spatial.Delaunay(pts).simplices, axis=1).astype(np.int32)
    return dfa.Mesh(pts, cells), rng


@pytest.mark.parametrize("gdim,n", [(2, 300), (3, 500)])
@pytest.mark.parametrize("scatter", [dfa.SCATTER_ATOMIC, dfa.SCATTER_BLOCKED, dfa.SCATTER_COLORED])
def test_p1_unstructured_vs_oracle(oracle, gdim, n, scatter):
    print(f"case unstructured p1 gdim={gdim} n={n} scatter={scatter}", flush=True)
    mesh, rng = _delaunay(gdim, n, seed=gdim * 100 + n)
    V = dfa.FunctionSpace(mesh, "P", 1)
    kappa, f = rng.uniform(0.5, 2, V.dim()), rng.uniform(-1, 1, V.dim())
    A = dfa.assemble(dfa.poisson_form(V, workloads.to_function(V, kappa)), scatter=scatter)
    b = dfa.assemble(dfa.load_form(V, workloads.to_function(V, f)), scatter=scatter)
    rp, col, vals = A.data()
    orp, ocol = oracle.sparsity(V.cell_dofs(), V.dim())
    np.testing.assert_array_equal(rp, orp)
    np.testing.assert_array_equal(col, ocol)
    ref = oracle.assemble_matrix(oracle.P1_POISSON, mesh.coordinates(), mesh.cells(),
                                 V.cell_dofs(), rp, col, c0=kappa)
    bref = oracle.assemble_vector(oracle.P1_LOAD, mesh.coordinates(), mesh.cells(),
                                  V.cell_dofs(), V.dim(), c0=f)
    assert _rel(vals, ref) < TOL
    assert _rel(b.get_local(), bref) < TOL
    # Dirichlet rows through SystemAssembler
    bc = dfa.DirichletBC(V, 0.0)
    g = rng.uniform(-1, 1, V.dim())
    bc = dfa.DirichletBC(V, g[bc.dofs], bc.dofs)
    As, bs = dfa.assemble_system(dfa.poisson_form(V, workloads.to_function(V, kappa)),
                                 dfa.load_form(V, workloads.to_function(V, f)), bc, scatter=scatter)
    mask = np.zeros(V.dim(), np.uint8)
    mask[bc.dofs] = 1
    gv = np.zeros(V.dim())
    gv[bc.dofs] = g[bc.dofs]
    ra, rb = oracle.assemble_system(oracle.P1_POISSON, oracle.P1_LOAD, mesh.coordinates(),
                                    mesh.cells(), V.cell_dofs(), rp, col, mask, gv, a_c0=kappa,
                                    L_c0=f)
    assert _rel(As.data()[2], ra) < TOL
    assert _rel(bs.get_local(), rb) < TOL
    # the boundary measure of the unit square / cube (python/test/unit/fem/test_assembler.py:140-153)
    one = dfa.Function(V, np.ones(V.dim()))
    assert abs(float(dfa.assemble(dfa.facet_load_form(V, one)).values.sum()) - 2.0 * gdim) < 1e-12
    # SpMV on the unstructured matrix
    S = sp.csr_matrix((vals, col, rp), shape=(V.dim(), V.dim()))
    xv = rng.uniform(-1, 1, V.dim())
    y = A.mult(torch.from_numpy(xv).cuda())
    torch.cuda.synchronize()
    assert _rel(y.cpu().numpy(), S @ xv) < 1e-14


@pytest.mark.parametrize("scatter", [dfa.SCATTER_ATOMIC, dfa.SCATTER_BLOCKED])
def test_vp2_unstructured_vs_oracle(oracle, scatter):
    print(f"case unstructured vp2 scatter={scatter}", flush=True)
    mesh, rng = _delaunay(3, 120, seed=7)
    V = dfa.VectorFunctionSpace(mesh, "P", 2)
    Q = V.coefficient_space()
    mu, lmbda = rng.uniform(0.5, 1.5, Q.dim()), rng.uniform(1, 10, Q.dim())
    f3 = np.array([0.0, 0.0, -1.0])
    A = dfa.assemble(dfa.elasticity_form(V, workloads.to_function(Q, mu),
                                         workloads.to_function(Q, lmbda)), scatter=scatter)
    b = dfa.assemble(dfa.body_force_form(V, f3), scatter=scatter)
    rp, col, vals = A.data()
    orp, ocol = oracle.sparsity(V.cell_dofs(), V.dim())
    np.testing.assert_array_equal(rp, orp)
    np.testing.assert_array_equal(col, ocol)
    ref = oracle.assemble_matrix(oracle.VP2_ELASTICITY, mesh.coordinates(), mesh.cells(),
                                 V.cell_dofs(), rp, col, coeff_dofs=Q.cell_dofs(), c0=mu, c1=lmbda)
    bref = oracle.assemble_vector(oracle.VP2_BODY_FORCE, mesh.coordinates(), mesh.cells(),
                                  V.cell_dofs(), V.dim(), f3=f3)
    assert _rel(vals, ref) < TOL
    assert _rel(b.get_local(), bref) < TOL


@pytest.mark.parametrize("gdim,n", [(2, 300), (3, 500)])
def test_p1_adjoint_unstructured_vs_oracle(oracle, gdim, n):
    """dJ/dkappa and dJ/dx on Delaunay meshes (in 3D through the edge-chain adjoint, whose
    chains also cover rows longer than 16 entries) against the oracle's closed forms."""
    from dolfin_amd import adjoint
    print(f"case unstructured adjoint gdim={gdim} n={n}", flush=True)
    mesh, rng = _delaunay(gdim, n, seed=gdim * 1000 + n)
    V = dfa.FunctionSpace(mesh, "P", 1)
    kappa = rng.uniform(0.5, 2.0, V.dim())
    f = rng.uniform(-1.0, 1.0, V.dim())
    u = rng.uniform(-1.0, 1.0, V.dim())
    kd, fd, ud = (torch.from_numpy(a).cuda() for a in (kappa, f, u))
    J, dk, _, dx = adjoint._run(V, dfa._lib.FORM_P1_POISSON, mesh.coordinates_device(), kd, None,
                                fd, None, ud)
    Jo, dko, dxo = oracle.adjoint_p1(mesh.coordinates(), mesh.cells(), V.cell_dofs(), kappa, f, u)
    assert abs(J.item() - Jo) < TOL * abs(Jo)
    assert _rel(dk.cpu().numpy(), dko) < TOL
    assert _rel(dx.cpu().numpy(), dxo) < TOL
