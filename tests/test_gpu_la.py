"""Device linear algebra on assembled matrices: CSR SpMV y = alpha A x + beta y and the diagonal
against scipy on the same CSR arrays (tolerance: relative L2 1e-14 for SpMV, bit-exact for the
diagonal), Jacobi CG against a direct solve (relative 1e-9) and the P1 nodal-exactness known
answer of -laplace(u) = -6, u = 1 + x^2 + 2 y^2. Vector-P2 matrices whose structure was
expanded from the node graph take the node-blocked SpMV kernel (both dof layouts); the
general streamed kernel runs for P1, for vector P2 with keep_diagonal (general structure
build) and under DFA_SPMV_CSR=1."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla
import torch

import dolfin_amd as dfa
from dolfin_amd import workloads

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _scipy(A):
    rp, col, v = A.data()
    n = rp.shape[0] - 1
    return sp.csr_matrix((v, col, rp), shape=(n, n))


def _p1_matrix(n, perm=None):
    mesh, V, kappa, f = workloads.p1_poisson_3d(n, seed=5)
    if perm is not None:  # renumbered dofs: "all" (random) or "swap" (two far-apart vertices)
        nv = mesh.num_vertices()
        p = np.random.default_rng(n).permutation(nv) if perm == "all" else np.arange(nv)
        if perm == "swap":
            p[[0, nv - 1]] = p[[nv - 1, 0]]
        V = dfa.FunctionSpace(mesh, "P", 1, node_permutation=p)
    return dfa.assemble(dfa.poisson_form(V, workloads.to_function(V, kappa)))


def _vp2_matrix(n, node_major=False, keep_diagonal=False):
    perm = None
    if node_major:
        m = dfa.UnitCubeMesh(n, n, n)
        perm = np.random.default_rng(n).permutation(m.num_vertices() + m.cell_edges()[1])
    mesh, V, mu, lmbda, f3 = workloads.vp2_elasticity_3d(n, seed=6, node_permutation=perm)
    Q = V.coefficient_space()
    a = dfa.elasticity_form(V, workloads.to_function(Q, mu), workloads.to_function(Q, lmbda))
    if keep_diagonal:
        asm = dfa.Assembler()
        asm.keep_diagonal = True
        A = dfa.Matrix()
        asm.assemble(A, a)
        return A
    return dfa.assemble(a)


# P1 row blocks read 16-bit column offsets when their columns span at most 65536 ("p1"), the
# 4-byte columns otherwise ("p1_all": randomly renumbered dofs, every block wide; "p1_swap":
# two far-apart vertices swapped, a few wide blocks among narrow ones)
@pytest.mark.parametrize("kind,n", [("p1", 7), ("p1", 24), ("p1_all", 40), ("p1_swap", 40), ("vp2", 3),
                                    ("vp2", 6), ("vp2_node_major", 4), ("vp2_node_major", 7),
                                    ("vp2_general", 4)])
def test_spmv_and_diagonal_vs_scipy(kind, n):
    print(f"case spmv {kind} n={n}", flush=True)
    if kind.startswith("p1"):
        A = _p1_matrix(n, perm=kind[3:] or None)
    else:
        A = _vp2_matrix(n, node_major=kind == "vp2_node_major", keep_diagonal=kind == "vp2_general")
    S = _scipy(A)
    rng = np.random.default_rng(n)
    x = rng.uniform(-1, 1, S.shape[0])
    y0 = rng.uniform(-1, 1, S.shape[0])
    xd = torch.from_numpy(x).cuda()
    y = A.mult(xd)
    torch.cuda.synchronize()
    assert _rel(y.cpu().numpy(), S @ x) < 1e-14
    yd = torch.from_numpy(y0.copy()).cuda()
    A.mult(xd, yd, alpha=-0.5, beta=2.0)
    torch.cuda.synchronize()
    assert _rel(yd.cpu().numpy(), 2.0 * y0 - 0.5 * (S @ x)) < 1e-14
    d = A.diagonal()
    torch.cuda.synchronize()
    assert np.array_equal(d.cpu().numpy(), S.diagonal())


@pytest.mark.parametrize("kind", ["vp2", "vp2_node_major", "p1", "p1_swap"])
def test_spmv_equals_plain_streamed_csr(monkeypatch, kind):
    # the node-blocked kernel (vector P2) and the 16-bit column offsets (P1) against the
    # streamed kernel on the 4-byte columns (DFA_SPMV_CSR=1): equal up to the summation order
    print(f"case spmv vs plain csr {kind}", flush=True)
    if kind.startswith("p1"):
        A = _p1_matrix(40, perm=kind[3:] or None)
    else:
        A = _vp2_matrix(5, node_major=kind == "vp2_node_major")
    x = torch.from_numpy(np.random.default_rng(11).uniform(-1, 1, A.structure()[0].numel() - 1)).cuda()
    y_node = A.mult(x)
    monkeypatch.setenv("DFA_SPMV_CSR", "1")
    y_csr = A.mult(x)
    torch.cuda.synchronize()
    assert (torch.linalg.norm(y_node - y_csr) / torch.linalg.norm(y_csr)).item() < 1e-14


def test_spmv_long_row_fan_mesh():
    # a fan of 5000 triangles around one vertex: its row holds 5001 nonzeros, more than one
    # streamed row block takes, and is summed by a whole workgroup
    m = 5000
    t = 2 * np.pi * np.arange(m) / m
    x = np.vstack([[0.0, 0.0], np.c_[np.cos(t), np.sin(t)]])
    cells = np.array([[0, 1 + i, 1 + (i + 1) % m] for i in range(m)], dtype=np.int32)
    cells.sort(axis=1)
    mesh = dfa.Mesh(x, cells)
    V = dfa.FunctionSpace(mesh, "P", 1)
    A = dfa.assemble(dfa.poisson_form(V, dfa.Function(V, np.ones(V.dim()))))
    S = _scipy(A)
    assert np.diff(S.indptr).max() == m + 1
    xv = np.random.default_rng(3).uniform(-1, 1, S.shape[0])
    y = A.mult(torch.from_numpy(xv).cuda())
    torch.cuda.synchronize()
    assert _rel(y.cpu().numpy(), S @ xv) < 1e-14
    assert np.array_equal(A.diagonal().cpu().numpy(), S.diagonal())


def test_spmv_accepts_vectors():
    # Vector objects are accepted wherever device tensors are
    A = _p1_matrix(3)
    S = _scipy(A)
    x = dfa.Vector(S.shape[0])
    x.values.fill_(1.0)
    y = dfa.Vector(S.shape[0])
    A.mult(x, y)
    torch.cuda.synchronize()
    np.testing.assert_allclose(y.get_local(), S @ np.ones(S.shape[0]), rtol=0, atol=1e-13)


@pytest.mark.parametrize("precond", ["jacobi", "none"])
def test_cg_vs_direct(precond):
    print(f"case cg {precond}", flush=True)
    mesh, V, kappa, f = workloads.p1_poisson_3d(10, seed=11)
    bc = dfa.DirichletBC(V, 0.5)
    A, b = dfa.assemble_system(dfa.poisson_form(V, workloads.to_function(V, kappa)),
                               dfa.load_form(V, workloads.to_function(V, f)), bc)
    x = dfa.Vector(V.dim())
    it = dfa.solve(A, x, b, "cg", precond, rtol=1e-12)
    ref = spla.spsolve(_scipy(A).tocsc(), b.get_local())
    assert it > 0
    assert _rel(x.get_local(), ref) < 1e-9


def test_cg_poisson_nodal_known_answer():
    # -laplace(u) = -6 with u = 1 + x^2 + 2 y^2 on the boundary: the P1 solution on a uniform
    # right-diagonal mesh is exact at the vertices
    mesh = dfa.UnitSquareMesh(16, 16)
    V = dfa.FunctionSpace(mesh, "P", 1)
    nv = mesh.num_vertices()
    vert_of_dof = np.empty(V.dim(), np.int64)
    vert_of_dof[V.cell_dofs().ravel()] = mesh.cells().ravel()
    xy = mesh.coordinates()[vert_of_dof]
    exact = 1 + xy[:, 0] ** 2 + 2 * xy[:, 1] ** 2
    bc = dfa.DirichletBC(V, 0.0)
    bc = dfa.DirichletBC(V, exact[bc.dofs], bc.dofs)
    one = dfa.Function(V, np.ones(V.dim()))
    f = dfa.Function(V, np.full(V.dim(), -6.0))
    A, b = dfa.assemble_system(dfa.poisson_form(V, one), dfa.load_form(V, f), bc)
    x = dfa.Vector(V.dim())
    dfa.solve(A, x, b, "cg", "jacobi", rtol=1e-13)
    assert nv == V.dim()
    assert np.max(np.abs(x.get_local() - exact)) < 1e-10


def test_cg_diagonal_system_stops_exactly():
    # every dof Dirichlet: A is the identity and b the boundary values. Jacobi-CG reaches r = 0
    # at iteration 1; later iterations must not divide 0 by 0 (p.Ap = 0), and the count is the
    # converged iteration, whatever the host check interval
    mesh, V, kappa, f = workloads.p1_poisson_3d(4, seed=2)
    g = np.random.default_rng(3).uniform(-1, 1, V.dim())
    bc = dfa.DirichletBC(V, g, dofs=np.arange(V.dim()))
    A, b = dfa.assemble_system(dfa.poisson_form(V, workloads.to_function(V, kappa)),
                               dfa.load_form(V, workloads.to_function(V, f)), bc)
    for every in (1, 10):
        x = dfa.Vector(V.dim())
        it = dfa.solve(A, x, b, "cg", "jacobi", rtol=1e-12, check_every=every)
        assert it == 1
        xv = x.get_local()
        assert np.all(np.isfinite(xv)) and np.max(np.abs(xv - g)) < 1e-15
    # the count does not depend on the check interval on a real system either
    bc2 = dfa.DirichletBC(V, 0.5)
    A2, b2 = dfa.assemble_system(dfa.poisson_form(V, workloads.to_function(V, kappa)),
                                 dfa.load_form(V, workloads.to_function(V, f)), bc2)
    its = set()
    sols = []
    for every in (1, 7, 10):
        x = dfa.Vector(V.dim())
        its.add(dfa.solve(A2, x, b2, "cg", "jacobi", rtol=1e-10, check_every=every))
        sols.append(x.get_local())
    assert len(its) == 1
    assert all(np.array_equal(sols[0], s) for s in sols[1:])


def _cg_reference(S, b, jacobi, rtol):
    """Preconditioned conjugate gradients in numpy, the algorithm the device solver runs."""
    d = S.diagonal()
    dinv = np.where(d != 0, 1.0 / np.where(d != 0, d, 1.0), 1.0) if jacobi else np.ones_like(b)
    x = np.zeros_like(b)
    r = b.copy()
    tol2 = (rtol * np.linalg.norm(b)) ** 2
    z = dinv * r
    p = z.copy()
    rz = r @ z
    for it in range(1, 10 * len(b) + 1):
        ap = S @ p
        alpha = rz / (p @ ap)
        x += alpha * p
        r -= alpha * ap
        z = dinv * r
        rz_new = r @ z
        if r @ r <= tol2:
            return x, it
        p = z + (rz_new / rz) * p
        rz = rz_new
    raise AssertionError("reference CG did not converge")


@pytest.mark.parametrize("precond", ["jacobi", "none"])
def test_cg_matches_reference_iterations(precond):
    # the device solver against the same iteration in numpy: the converged iteration agrees up
    # to the summation order of the reductions (at most one step), and so does the solution
    print(f"case cg reference {precond}", flush=True)
    mesh, V, kappa, f = workloads.p1_poisson_3d(9, seed=21)
    bc = dfa.DirichletBC(V, 0.25)
    A, b = dfa.assemble_system(dfa.poisson_form(V, workloads.to_function(V, kappa)),
                               dfa.load_form(V, workloads.to_function(V, f)), bc)
    x = dfa.Vector(V.dim())
    it = dfa.solve(A, x, b, "cg", precond, rtol=1e-11)
    xr, itr = _cg_reference(_scipy(A), b.get_local(), precond == "jacobi", 1e-11)
    assert abs(it - itr) <= 1
    assert _rel(x.get_local(), xr) < 1e-9


def test_cg_vector_p2_elasticity():
    # elasticity with the boundary clamped: the node-blocked SpMV inside the solver
    print("case cg vector P2", flush=True)
    mesh, V, mu, lmbda, f3 = workloads.vp2_elasticity_3d(3, seed=4)
    Q = V.coefficient_space()
    a = dfa.elasticity_form(V, workloads.to_function(Q, mu), workloads.to_function(Q, lmbda))
    A, b = dfa.assemble_system(a, dfa.body_force_form(V, f3), dfa.DirichletBC(V, 0.0))
    x = dfa.Vector(V.dim())
    it = dfa.solve(A, x, b, "cg", "jacobi", rtol=1e-12)
    ref = spla.spsolve(_scipy(A).tocsc(), b.get_local())
    assert it > 0
    assert _rel(x.get_local(), ref) < 1e-9


def test_cg_not_converged_raises():
    mesh, V, kappa, f = workloads.p1_poisson_3d(6, seed=22)
    A, b = dfa.assemble_system(dfa.poisson_form(V, workloads.to_function(V, kappa)),
                               dfa.load_form(V, workloads.to_function(V, f)), dfa.DirichletBC(V, 0.0))
    x = dfa.Vector(V.dim())
    with pytest.raises(RuntimeError, match="did not converge in 3 iterations"):
        dfa.solve(A, x, b, "cg", "jacobi", rtol=1e-14, maxiter=3)
