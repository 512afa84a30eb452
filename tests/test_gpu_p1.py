"""GPU parity of P1 assembly (HIP kernels through the C ABI) against the CPU oracle.

Bar: CSR structure and dof indexing bit-exact; values within relative Frobenius / L2 error
1e-12 in FP64 (summation order differs: atomics and LDS accumulation vs ascending cells).
"""
import numpy as np
import pytest
import torch

import dolfin_amd as dfa
from dolfin_amd import workloads

pytestmark = pytest.mark.gpu

TOL = 1e-12
SCATTERS = [dfa.SCATTER_ATOMIC, dfa.SCATTER_BLOCKED, dfa.SCATTER_COLORED]


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _check_structure(oracle, A, V):
    rp, col, vals = A.data()
    orp, ocol = oracle.sparsity(V.cell_dofs(), V.dim())
    np.testing.assert_array_equal(rp, orp)
    np.testing.assert_array_equal(col, ocol)
    return rp, col, vals


@pytest.mark.parametrize("scatter", SCATTERS)
def test_unit_square_64_right(oracle, scatter):
    mesh, V, kappa, f = workloads.p1_poisson_2d(64)
    print(f"case unit_square_64 scatter={scatter}", flush=True)
    K = workloads.to_function(V, kappa)
    F = workloads.to_function(V, f)
    A = dfa.assemble(dfa.poisson_form(V, K), scatter=scatter)
    b = dfa.assemble(dfa.load_form(V, F), scatter=scatter)
    rp, col, vals = _check_structure(oracle, A, V)
    ref = oracle.assemble_matrix(oracle.P1_POISSON, mesh.coordinates(), mesh.cells(), V.cell_dofs(),
                                 rp, col, c0=kappa)
    bref = oracle.assemble_vector(oracle.P1_LOAD, mesh.coordinates(), mesh.cells(), V.cell_dofs(),
                                  V.dim(), c0=f)
    assert rel(vals, ref) < TOL
    assert rel(b.get_local(), bref) < TOL
    assert col.shape[0] == 29057


@pytest.mark.parametrize("scatter", SCATTERS)
@pytest.mark.parametrize("n", [3, 8, 17])
def test_unit_cube_jittered(oracle, scatter, n):
    print(f"case unit_cube n={n} scatter={scatter}", flush=True)
    mesh, V, kappa, f = workloads.p1_poisson_3d(n, seed=n)
    K = workloads.to_function(V, kappa)
    F = workloads.to_function(V, f)
    A = dfa.assemble(dfa.poisson_form(V, K), scatter=scatter)
    M = dfa.assemble(dfa.mass_form(V), scatter=scatter)
    b = dfa.assemble(dfa.load_form(V, F), scatter=scatter)
    rp, col, vals = _check_structure(oracle, A, V)
    x, c, cd = mesh.coordinates(), mesh.cells(), V.cell_dofs()
    assert rel(vals, oracle.assemble_matrix(oracle.P1_POISSON, x, c, cd, rp, col, c0=kappa)) < TOL
    assert rel(M.data()[2], oracle.assemble_matrix(oracle.P1_MASS, x, c, cd, rp, col)) < TOL
    assert rel(b.get_local(), oracle.assemble_vector(oracle.P1_LOAD, x, c, cd, V.dim(), c0=f)) < TOL


@pytest.mark.parametrize("scatter", SCATTERS)
def test_interval_known_answer(scatter):
    # python/test/unit/fem/test_assembler.py:39-55
    mesh = dfa.UnitIntervalMesh(48)
    V = dfa.FunctionSpace(mesh, "P", 1)
    one = dfa.Function(V, torch.ones(V.dim(), dtype=torch.float64))
    ten = dfa.Function(V, torch.full((V.dim(),), 10.0, dtype=torch.float64))
    A = dfa.assemble(dfa.poisson_form(V, one), scatter=scatter)
    b = dfa.assemble(dfa.load_form(V, ten), scatter=scatter)
    assert round(A.norm("frobenius") - 811.75365721381274397572, 10) == 0
    assert round(b.norm("l2") - 1.43583841167606474087, 10) == 0


@pytest.mark.parametrize("scatter", SCATTERS)
def test_system_assembler_known_answer(scatter):
    # python/test/unit/fem/test_system_assembler.py:73-118
    mesh = dfa.UnitCubeMesh(4, 4, 4)
    V = dfa.FunctionSpace(mesh, "P", 1)
    one = dfa.Function(V, torch.ones(V.dim(), dtype=torch.float64))
    ten = dfa.Function(V, torch.full((V.dim(),), 10.0, dtype=torch.float64))
    bc = dfa.DirichletBC(V, 1.0, "on_boundary")
    a, L = dfa.poisson_form(V, one), dfa.load_form(V, ten)
    A, b = dfa.assemble_system(a, L, bc, scatter=scatter)
    assert round(A.norm("frobenius") - 96.847818767384, 10) == 0
    assert round(b.norm("l2") - 96.564760289080, 10) == 0
    asm = dfa.SystemAssembler(a, L, bc, scatter=scatter)
    A2 = dfa.Matrix()
    asm.assemble(A2)
    assert round(A2.norm("frobenius") - 96.847818767384, 10) == 0
    asm.assemble(A2)  # second time: re-assembly zeroes first
    assert round(A2.norm("frobenius") - 96.847818767384, 10) == 0
    b2 = dfa.Vector()
    asm.assemble(b=b2)
    assert round(b2.norm("l2") - 96.564760289080, 10) == 0
    asm.assemble(b=b2)
    assert round(b2.norm("l2") - 96.564760289080, 10) == 0


@pytest.mark.parametrize("scatter", SCATTERS)
def test_system_assembler_vs_oracle(oracle, scatter):
    mesh, V, kappa, f = workloads.p1_poisson_3d(6, seed=4)
    rng = np.random.default_rng(9)
    gvals = rng.uniform(-1, 1, V.dim())
    bc = dfa.DirichletBC(V, 0.0)
    bc_vals = np.zeros(V.dim())
    bc_vals[bc.dofs] = gvals[bc.dofs]
    bc = dfa.DirichletBC(V, gvals[bc.dofs], bc.dofs)
    A, b = dfa.assemble_system(dfa.poisson_form(V, workloads.to_function(V, kappa)),
                               dfa.load_form(V, workloads.to_function(V, f)), bc, scatter=scatter)
    rp, col, vals = _check_structure(oracle, A, V)
    mask = np.zeros(V.dim(), np.uint8)
    mask[bc.dofs] = 1
    ra, rb = oracle.assemble_system(oracle.P1_POISSON, oracle.P1_LOAD, mesh.coordinates(),
                                    mesh.cells(), V.cell_dofs(), rp, col, mask, bc_vals,
                                    a_c0=kappa, L_c0=f)
    assert rel(vals, ra) < TOL
    assert rel(b.get_local(), rb) < TOL


@pytest.mark.parametrize("scatter", SCATTERS)
def test_reassembly_and_add_values(oracle, scatter):
    mesh, V, kappa, f = workloads.p1_poisson_3d(5, seed=1)
    K = workloads.to_function(V, kappa)
    form = dfa.poisson_form(V, K)
    A = dfa.assemble(form, scatter=scatter)
    v1 = A.values.clone()
    dfa.assemble(form, tensor=A, scatter=scatter)  # zero + reassemble
    # same tensor up to the summation order of the accumulation
    assert rel(A.values.cpu().numpy(), v1.cpu().numpy()) < 1e-14
    dfa.assemble(form, tensor=A, add_values=True, scatter=scatter)
    assert rel(A.values.cpu().numpy(), 2 * v1.cpu().numpy()) < TOL


def test_keep_diagonal_and_permuted_dofs(oracle):
    rng = np.random.default_rng(5)
    mesh = dfa.UnitCubeMesh(4, 3, 5)
    perm = rng.permutation(mesh.num_vertices())
    V = dfa.FunctionSpace(mesh, "P", 1, node_permutation=perm)
    kappa = rng.uniform(0.5, 2, V.dim())
    for scatter in SCATTERS:
        A = dfa.assemble(dfa.poisson_form(V, workloads.to_function(V, kappa)), scatter=scatter,
                         keep_diagonal=True)
        rp, col, vals = _check_structure(oracle, A, V)
        # kappa lives on the permuted space too: coefficient dof = perm[vertex]
        ref = oracle.assemble_matrix(oracle.P1_POISSON, mesh.coordinates(), mesh.cells(),
                                     V.cell_dofs(), rp, col, coeff_dofs=V.cell_dofs(), c0=kappa)
        assert rel(vals, ref) < TOL


@pytest.mark.parametrize("scatter", SCATTERS)
def test_rcm_reordered_system_vs_oracle(oracle, scatter):
    # DOLFIN's default numbering (reorder_dofs_serial): reverse Cuthill-McKee node order
    rng = np.random.default_rng(12)
    mesh = dfa.UnitCubeMesh(8, 7, 6)
    dfa.jitter_interior(mesh, 0.15 / 8, rng)
    V = dfa.FunctionSpace(mesh, "P", 1, reorder=True)
    kappa, f = rng.uniform(0.5, 2, V.dim()), rng.uniform(-1, 1, V.dim())
    bc = dfa.DirichletBC(V, 0.0)
    g = rng.uniform(-1, 1, V.dim())
    bc = dfa.DirichletBC(V, g[bc.dofs], bc.dofs)
    A, b = dfa.assemble_system(dfa.poisson_form(V, workloads.to_function(V, kappa)),
                               dfa.load_form(V, workloads.to_function(V, f)), bc, scatter=scatter)
    rp, col, vals = _check_structure(oracle, A, V)
    mask = np.zeros(V.dim(), np.uint8)
    mask[bc.dofs] = 1
    gv = np.zeros(V.dim())
    gv[bc.dofs] = g[bc.dofs]
    ra, rb = oracle.assemble_system(oracle.P1_POISSON, oracle.P1_LOAD, mesh.coordinates(),
                                    mesh.cells(), V.cell_dofs(), rp, col, mask, gv,
                                    coeff_dofs=V.cell_dofs(), a_c0=kappa, L_c0=f)
    assert rel(vals, ra) < TOL
    assert rel(b.get_local(), rb) < TOL


class _PaddedP1:
    """A P1 space with `extra` trailing dofs that belong to no cell (empty CSR rows)."""

    element = dfa._lib.ELEMENT_P1

    def __init__(self, mesh, extra):
        self._mesh, self._extra = mesh, extra
        self._W = dfa.FunctionSpace(mesh, "P", 1)

    def mesh(self):
        return self._mesh

    def cell_dofs(self):
        return self._mesh.cells()

    def dim(self):
        return self._mesh.num_vertices() + self._extra

    def coefficient_space(self):
        return self._W


@pytest.mark.parametrize("keep_diagonal", [False, True])
@pytest.mark.parametrize("where", ["device", "host"])
def test_csr_structure_build_edge_rows(oracle, keep_diagonal, where, monkeypatch):
    # the structure built on the device (radix sort of the cell pairs) and on the host give
    # the oracle's sparsity, including rows in no cell (empty, or the diagonal alone)
    if where == "host":
        monkeypatch.setenv("DFA_HOST_CSR", "1")
    mesh = dfa.UnitCubeMesh(3, 2, 2)
    V = _PaddedP1(mesh, 5)
    eng = dfa.fem.Engine(V, keep_diagonal=keep_diagonal)
    rp, col = eng.structure()
    torch.cuda.synchronize()
    orp, ocol = oracle.sparsity(mesh.cells(), V.dim(), keep_diagonal=keep_diagonal)
    np.testing.assert_array_equal(rp.cpu().numpy(), orp)
    np.testing.assert_array_equal(col.cpu().numpy(), ocol)


def test_csr_device_build_equals_host_vp2(monkeypatch):
    mesh, V, mu, lmbda, f3 = workloads.vp2_elasticity_3d(5, seed=2)
    dev = dfa.fem.Engine(V)
    monkeypatch.setenv("DFA_HOST_CSR", "1")
    host = dfa.fem.Engine(V)
    a, b = dev.structure(), host.structure()
    torch.cuda.synchronize()
    assert dev.nnz == host.nnz
    assert torch.equal(a[0], b[0]) and torch.equal(a[1], b[1])


@pytest.mark.parametrize("layout", ["component_major", "node_major"])
def test_csr_vp2_node_blocked_build(oracle, layout, monkeypatch):
    # The vector-P2 device build sorts the node graph and expands every node entry to its
    # 3 x 3 block when the dofs are node-blocked (component-major UFC numbering, or node-major
    # after a node permutation). It must equal the oracle's sparsity and the general device
    # build (forced by DFA_CSR_GENERAL, and taken by keep_diagonal, which adds nothing here:
    # every vector-P2 dof row holds its diagonal).
    rng = np.random.default_rng(7)
    mesh = dfa.UnitCubeMesh(3, 2, 4)
    perm = None if layout == "component_major" else rng.permutation(
        mesh.num_vertices() + mesh.cell_edges()[1]).astype(np.int32)
    V = dfa.VectorFunctionSpace(mesh, "P", 2, node_permutation=perm)
    fast = dfa.fem.Engine(V)
    diag = dfa.fem.Engine(V, keep_diagonal=True)
    monkeypatch.setenv("DFA_CSR_GENERAL", "1")
    general = dfa.fem.Engine(V)
    a, b, c = fast.structure(), general.structure(), diag.structure()
    torch.cuda.synchronize()
    orp, ocol = oracle.sparsity(V.cell_dofs(), V.dim())
    np.testing.assert_array_equal(a[0].cpu().numpy(), orp)
    np.testing.assert_array_equal(a[1].cpu().numpy(), ocol)
    assert torch.equal(a[0], b[0]) and torch.equal(a[1], b[1])
    assert torch.equal(a[0], c[0]) and torch.equal(a[1], c[1])


def test_blocked_equals_atomic_128_properties():
    """Full-size P1 128^3 (12.58M cells): size-independent properties -- blocked and atomic
    scatters agree, Poisson rows sum to zero (K 1 = 0), K is symmetric, sum(b) = int f."""
    mesh, V, kappa, f = workloads.p1_poisson_3d(128)
    print("case p1_128 properties", flush=True)
    K = workloads.to_function(V, kappa)
    F = workloads.to_function(V, f)
    Ab = dfa.assemble(dfa.poisson_form(V, K), scatter=dfa.SCATTER_BLOCKED)
    Aa = dfa.assemble(dfa.poisson_form(V, K), scatter=dfa.SCATTER_ATOMIC)
    va, vb = Aa.values, Ab.values
    assert (torch.linalg.norm(va - vb) / torch.linalg.norm(va)).item() < TOL
    rp, col = Ab.structure()
    rows = torch.repeat_interleave(torch.arange(V.dim(), device="cuda"),
                                   (rp[1:] - rp[:-1]).long())
    rowsum = torch.zeros(V.dim(), dtype=torch.float64, device="cuda").index_add_(0, rows, vb)
    diag = vb[rows == col.long()]
    assert (rowsum.abs().max() / diag.abs().max()).item() < 1e-13
    # symmetry: compare (r, c) with (c, r) through a key sort
    n = V.dim()
    key = rows * n + col.long()
    tkey = col.long() * n + rows
    perm = torch.argsort(tkey)
    assert torch.equal(key, tkey[perm])
    assert (torch.linalg.norm(vb - vb[perm]) / torch.linalg.norm(vb)).item() < 1e-15
    bb = dfa.assemble(dfa.load_form(V, F), scatter=dfa.SCATTER_BLOCKED).values
    ba = dfa.assemble(dfa.load_form(V, F), scatter=dfa.SCATTER_ATOMIC).values
    assert (torch.linalg.norm(ba - bb) / torch.linalg.norm(ba)).item() < TOL
    # graph-coloured scatter: same values; its colours separate every pair of cells that
    # share a dof (checked from the colour offsets and the colour-order dofs on the host)
    Ac = dfa.assemble(dfa.poisson_form(V, K), scatter=dfa.SCATTER_COLORED)
    assert (torch.linalg.norm(Ac.values - vb) / torch.linalg.norm(vb)).item() < TOL
    bc = dfa.assemble(dfa.load_form(V, F), scatter=dfa.SCATTER_COLORED).values
    assert (torch.linalg.norm(bc - bb) / torch.linalg.norm(bb)).item() < TOL
    ptr = dfa.engine_for(V).color_ptr()
    assert ptr[0] == 0 and ptr[-1] == mesh.num_cells() and np.all(np.diff(ptr) > 0)
    assert 24 <= len(ptr) - 1 <= 64  # at least the 24 cells around an interior vertex


def test_scatter_autotuner_picks_a_strategy():
    """autotune.tune_scatter times every P1 strategy on the forms and returns the fastest;
    the three strategies give the same matrix."""
    from dolfin_amd import autotune
    mesh, V, kappa, f = workloads.p1_poisson_3d(16, seed=3)
    print("case scatter autotuner n=16", flush=True)
    a = dfa.poisson_form(V, workloads.to_function(V, kappa))
    best, times = autotune.tune_scatter([a], reps=2)
    assert set(times) == set(SCATTERS) and best in SCATTERS
    assert times[best] == min(times.values()) and all(t > 0 for t in times.values())
    ref = dfa.assemble(a, scatter=dfa.SCATTER_ATOMIC).values
    for sc in SCATTERS:
        v = dfa.assemble(a, scatter=sc).values
        assert (torch.linalg.norm(v - ref) / torch.linalg.norm(ref)).item() < TOL
