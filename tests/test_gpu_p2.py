"""GPU parity of vector-P2 elasticity assembly against the CPU oracle (structure bit-exact,
values <= 1e-12 relative), plus full-size properties (rigid-body null space)."""
import numpy as np
import pytest
import torch

import dolfin_amd as dfa
from dolfin_amd import workloads

pytestmark = pytest.mark.gpu

TOL = 1e-12


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _forms(V, mu, lmbda, f3):
    Q = V.coefficient_space()
    a = dfa.elasticity_form(V, workloads.to_function(Q, mu), workloads.to_function(Q, lmbda))
    L = dfa.body_force_form(V, f3)
    return a, L


@pytest.mark.parametrize("scatter", [dfa.SCATTER_ATOMIC, dfa.SCATTER_BLOCKED])
@pytest.mark.parametrize("n,block", [(2, False), (3, False), (3, True), (5, False)])
def test_vp2_vs_oracle(oracle, n, block, scatter):
    print(f"case vp2 n={n} block={block} scatter={scatter}", flush=True)
    perm = None
    if block:
        m = dfa.UnitCubeMesh(n, n, n)
        perm = np.random.default_rng(n).permutation(m.num_vertices() + m.cell_edges()[1])
    mesh, V, mu, lmbda, f3 = workloads.vp2_elasticity_3d(n, seed=n, node_permutation=perm)
    a, L = _forms(V, mu, lmbda, f3)
    A = dfa.assemble(a, scatter=scatter)
    b = dfa.assemble(L, scatter=scatter)
    rp, col, vals = A.data()
    orp, ocol = oracle.sparsity(V.cell_dofs(), V.dim())
    np.testing.assert_array_equal(rp, orp)
    np.testing.assert_array_equal(col, ocol)
    kd = V.coefficient_space().cell_dofs()
    ref = oracle.assemble_matrix(oracle.VP2_ELASTICITY, mesh.coordinates(), mesh.cells(),
                                 V.cell_dofs(), rp, col, coeff_dofs=kd, c0=mu, c1=lmbda)
    bref = oracle.assemble_vector(oracle.VP2_BODY_FORCE, mesh.coordinates(), mesh.cells(),
                                  V.cell_dofs(), V.dim(), f3=f3)
    assert rel(vals, ref) < TOL
    assert rel(b.get_local(), bref) < TOL


@pytest.mark.parametrize("scatter", [dfa.SCATTER_ATOMIC, dfa.SCATTER_BLOCKED])
def test_vp2_rcm_reordered_vs_oracle(oracle, scatter):
    # a reverse Cuthill-McKee node order, dofs 3 * node + component (the block-interleaved
    # layout DOLFIN uses after reordering, DofMapBuilder.cpp:1689)
    m = dfa.UnitCubeMesh(4, 4, 4)
    ce, ne = m.cell_edges()
    nodes = np.concatenate([m.cells(), m.num_vertices() + ce], axis=1)
    perm = dfa.node_reordering(nodes, m.num_vertices() + ne)
    mesh, V, mu, lmbda, f3 = workloads.vp2_elasticity_3d(4, seed=9, node_permutation=perm)
    a, L = _forms(V, mu, lmbda, f3)
    A = dfa.assemble(a, scatter=scatter)
    b = dfa.assemble(L, scatter=scatter)
    rp, col, vals = A.data()
    orp, ocol = oracle.sparsity(V.cell_dofs(), V.dim())
    np.testing.assert_array_equal(rp, orp)
    np.testing.assert_array_equal(col, ocol)
    kd = V.coefficient_space().cell_dofs()
    ref = oracle.assemble_matrix(oracle.VP2_ELASTICITY, mesh.coordinates(), mesh.cells(),
                                 V.cell_dofs(), rp, col, coeff_dofs=kd, c0=mu, c1=lmbda)
    bref = oracle.assemble_vector(oracle.VP2_BODY_FORCE, mesh.coordinates(), mesh.cells(),
                                  V.cell_dofs(), V.dim(), f3=f3)
    assert rel(vals, ref) < TOL
    assert rel(b.get_local(), bref) < TOL


@pytest.mark.parametrize("scatter", [dfa.SCATTER_ATOMIC, dfa.SCATTER_BLOCKED])
def test_vp2_system_bc_vs_oracle(oracle, scatter):
    mesh, V, mu, lmbda, f3 = workloads.vp2_elasticity_3d(3, seed=8)
    a, L = _forms(V, mu, lmbda, f3)
    bc = dfa.DirichletBC(V, 0.25)
    A, b = dfa.assemble_system(a, L, bc, scatter=scatter)
    rp, col, vals = A.data()
    mask = np.zeros(V.dim(), np.uint8)
    mask[bc.dofs] = 1
    g = np.where(mask > 0, 0.25, 0.0)
    kd = V.coefficient_space().cell_dofs()
    ra, rb = oracle.assemble_system(oracle.VP2_ELASTICITY, oracle.VP2_BODY_FORCE,
                                    mesh.coordinates(), mesh.cells(), V.cell_dofs(), rp, col, mask,
                                    g, coeff_dofs=kd, a_c0=mu, a_c1=lmbda, f3=f3)
    assert rel(vals, ra) < TOL
    assert rel(b.get_local(), rb) < TOL


def test_vp2_32_rigid_modes():
    """32^3 (196,608 cells, 823,875 dofs): K r = 0 for the six rigid-body modes, which P2
    interpolates exactly, and sum of the z-load = -volume."""
    mesh, V, mu, lmbda, f3 = workloads.vp2_elasticity_3d(32)
    print("case vp2_32 rigid modes", flush=True)
    a, L = _forms(V, mu, lmbda, f3)
    A = dfa.assemble(a)
    Aa = dfa.assemble(a, scatter=dfa.SCATTER_ATOMIC)
    assert (torch.linalg.norm(A.values - Aa.values) / torch.linalg.norm(Aa.values)).item() < TOL
    rp, col = A.structure()
    rows = torch.repeat_interleave(torch.arange(V.dim(), device="cuda"), (rp[1:] - rp[:-1]).long())
    nv = mesh.num_vertices()
    ce, ne = mesh.cell_edges()
    x = mesh.coordinates()
    # node coordinates: vertices then edge midpoints
    loc = np.array([[2, 3], [1, 3], [1, 2], [0, 3], [0, 2], [0, 1]])
    ex = np.zeros((ne, 3))
    c = mesh.cells()
    ex[ce.ravel()] = (0.5 * (x[c[:, loc[:, 0]]] + x[c[:, loc[:, 1]]])).reshape(-1, 3)
    X = np.vstack([x, ex])
    nn = nv + ne
    scale = A.values.abs().max().item()
    modes = [np.eye(3)[k] * np.ones((nn, 3)) for k in range(3)]
    modes += [np.cross(axis, X - 0.5) for axis in np.eye(3)]
    for m in modes:
        r = torch.from_numpy(m.T.ravel().copy()).cuda()  # component-major dof = comp*nn + node
        Kr = torch.zeros(V.dim(), dtype=torch.float64, device="cuda").index_add_(
            0, rows, A.values * r[col.long()])
        assert Kr.abs().max().item() < 1e-11 * scale * max(1.0, float(np.abs(m).max()))
    b = dfa.assemble(L).values
    assert abs(b.sum().item() + 1.0) < 1e-12
