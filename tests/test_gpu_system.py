"""GPU parity of the kernels the benchmark times and of the SystemAssembler surface.

The benchmark's passes are SystemAssembler::assemble(A, b) without Dirichlet conditions:
P1 kappa-Poisson + load on tetrahedra runs the fused edge-chain kernel, vector-P2
elasticity + body force the fused node-block kernel. These tests run exactly those
instantiations (the library names the kernel a pass launches, and each test asserts it)
against the oracle's SystemAssembler loop (dolfin/fem/SystemAssembler.cpp:303-492), with
vertex and permuted dofs, and with add_values (AssemblerBase.h:53). Also: the incremental
assemble(A, b, x0) (SystemAssembler.cpp:252-275), add_values on the vector-P2 blocked path,
and host-side add_local blocks (GenericMatrix.h:106-112).

Bar: CSR structure bit-exact; values within relative Frobenius / L2 error 1e-12 (FP64, the
summation order differs).
"""
import numpy as np
import pytest
import torch

import dolfin_amd as dfa
from dolfin_amd import _lib, workloads

pytestmark = pytest.mark.gpu

TOL = 1e-12


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _p1_case(n, permuted, seed):
    mesh, V0, kappa, f = workloads.p1_poisson_3d(n, seed=seed)
    if not permuted:
        return mesh, V0, kappa, f
    perm = np.random.default_rng(seed + 100).permutation(mesh.num_vertices())
    V = dfa.FunctionSpace(mesh, "P", 1, node_permutation=perm)
    # the same nodal values, now stored at the permuted dofs
    kp, fp = np.empty_like(kappa), np.empty_like(f)
    kp[perm], fp[perm] = kappa, f
    return mesh, V, kp, fp


@pytest.mark.parametrize("permuted", [False, True])
@pytest.mark.parametrize("n", [3, 8, 17])
def test_fused_p1_system_no_bc(oracle, n, permuted):
    print(f"case fused p1 system n={n} permuted={permuted}", flush=True)
    mesh, V, kappa, f = _p1_case(n, permuted, seed=30 + n)
    eng = dfa.engine_for(V)
    kname = eng.pass_kernel(_lib.FORM_P1_POISSON, _lib.FORM_P1_LOAD)
    # the benchmark's kernel: fused matrix + vector edge-chain walk, 512 threads
    assert kname.startswith("p1_chain_kernel<WANT_B=1,T=512,"), kname
    a = dfa.poisson_form(V, workloads.to_function(V, kappa))
    L = dfa.load_form(V, workloads.to_function(V, f))
    A, b = dfa.assemble_system(a, L)
    rp, col, vals = A.data()
    orp, ocol = oracle.sparsity(V.cell_dofs(), V.dim())
    np.testing.assert_array_equal(rp, orp)
    np.testing.assert_array_equal(col, ocol)
    nomask = np.zeros(V.dim(), np.uint8)
    ra, rb = oracle.assemble_system(oracle.P1_POISSON, oracle.P1_LOAD, mesh.coordinates(),
                                    mesh.cells(), V.cell_dofs(), rp, col, nomask,
                                    np.zeros(V.dim()), coeff_dofs=V.cell_dofs(), a_c0=kappa,
                                    L_c0=f)
    assert rel(vals, ra) < TOL
    assert rel(b.get_local(), rb) < TOL
    # add_values: the same pass adds onto what the tensors hold
    dfa.assemble_system(a, L, add_values=True, A_tensor=A, b_tensor=b)
    assert rel(A.values.cpu().numpy(), 2 * ra) < TOL
    assert rel(b.get_local(), 2 * rb) < TOL
    # re-assembly without add_values overwrites
    dfa.assemble_system(a, L, A_tensor=A, b_tensor=b)
    assert rel(A.values.cpu().numpy(), ra) < TOL
    assert rel(b.get_local(), rb) < TOL


def _vp2_case(n, permuted, seed):
    perm = None
    if permuted:
        m = dfa.UnitCubeMesh(n, n, n)
        perm = np.random.default_rng(seed).permutation(m.num_vertices() + m.cell_edges()[1])
    return workloads.vp2_elasticity_3d(n, seed=seed, node_permutation=perm)


@pytest.mark.parametrize("n,permuted", [(2, False), (3, False), (3, True), (5, False)])
def test_fused_vp2_system_no_bc(oracle, n, permuted):
    print(f"case fused vp2 system n={n} permuted={permuted}", flush=True)
    mesh, V, mu, lmbda, f3 = _vp2_case(n, permuted, seed=40 + n)
    eng = dfa.engine_for(V)
    kname = eng.pass_kernel(_lib.FORM_VP2_ELASTICITY, _lib.FORM_VP2_BODY_FORCE)
    assert kname == "vp2_blocked_kernel<WANT_A=1,WANT_B=1,BC=0>", kname
    Q = V.coefficient_space()
    a = dfa.elasticity_form(V, workloads.to_function(Q, mu), workloads.to_function(Q, lmbda))
    L = dfa.body_force_form(V, f3)
    A, b = dfa.assemble_system(a, L)
    rp, col, vals = A.data()
    orp, ocol = oracle.sparsity(V.cell_dofs(), V.dim())
    np.testing.assert_array_equal(rp, orp)
    np.testing.assert_array_equal(col, ocol)
    kd = Q.cell_dofs()
    ra, rb = oracle.assemble_system(oracle.VP2_ELASTICITY, oracle.VP2_BODY_FORCE,
                                    mesh.coordinates(), mesh.cells(), V.cell_dofs(), rp, col,
                                    np.zeros(V.dim(), np.uint8), np.zeros(V.dim()),
                                    coeff_dofs=kd, a_c0=mu, a_c1=lmbda, f3=f3)
    assert rel(vals, ra) < TOL
    assert rel(b.get_local(), rb) < TOL
    dfa.assemble_system(a, L, add_values=True, A_tensor=A, b_tensor=b)
    assert rel(A.values.cpu().numpy(), 2 * ra) < TOL
    assert rel(b.get_local(), 2 * rb) < TOL


@pytest.mark.parametrize("scatter", [dfa.SCATTER_BLOCKED, dfa.SCATTER_ATOMIC])
def test_vp2_assemble_add_values(oracle, scatter):
    # Assembler::assemble with add_values on every vector-P2 scatter (AssemblerBase.h:53)
    mesh, V, mu, lmbda, f3 = workloads.vp2_elasticity_3d(3, seed=7)
    Q = V.coefficient_space()
    a = dfa.elasticity_form(V, workloads.to_function(Q, mu), workloads.to_function(Q, lmbda))
    L = dfa.body_force_form(V, f3)
    A = dfa.assemble(a, scatter=scatter)
    b = dfa.assemble(L, scatter=scatter)
    rp, col, vals = A.data()
    ref = oracle.assemble_matrix(oracle.VP2_ELASTICITY, mesh.coordinates(), mesh.cells(),
                                 V.cell_dofs(), rp, col, coeff_dofs=Q.cell_dofs(), c0=mu, c1=lmbda)
    bref = oracle.assemble_vector(oracle.VP2_BODY_FORCE, mesh.coordinates(), mesh.cells(),
                                  V.cell_dofs(), V.dim(), f3=f3)
    assert rel(vals, ref) < TOL
    dfa.assemble(a, tensor=A, add_values=True, scatter=scatter)
    dfa.assemble(L, tensor=b, add_values=True, scatter=scatter)
    assert rel(A.values.cpu().numpy(), 2 * ref) < TOL
    assert rel(b.get_local(), 2 * bref) < TOL


@pytest.mark.parametrize("scatter", [dfa.SCATTER_BLOCKED, dfa.SCATTER_ATOMIC, dfa.SCATTER_COLORED])
def test_p1_system_x0_vs_oracle(oracle, scatter):
    # SystemAssembler::assemble(A, b, x0) and assemble(b, x0) (SystemAssembler.h:79,84)
    mesh, V, kappa, f = workloads.p1_poisson_3d(5, seed=12)
    rng = np.random.default_rng(13)
    g = rng.uniform(-1, 1, V.dim())
    x0 = rng.uniform(-1, 1, V.dim())
    bc0 = dfa.DirichletBC(V, 0.0)
    bc = dfa.DirichletBC(V, g[bc0.dofs], bc0.dofs)
    a = dfa.poisson_form(V, workloads.to_function(V, kappa))
    L = dfa.load_form(V, workloads.to_function(V, f))
    x0t = torch.from_numpy(x0).cuda()
    A, b = dfa.assemble_system(a, L, bc, x0=x0t, scatter=scatter)
    rp, col, vals = A.data()
    mask = np.zeros(V.dim(), np.uint8)
    mask[bc.dofs] = 1
    gv = np.where(mask > 0, g, 0.0)
    ra, rb = oracle.assemble_system(oracle.P1_POISSON, oracle.P1_LOAD, mesh.coordinates(),
                                    mesh.cells(), V.cell_dofs(), rp, col, mask, gv, a_c0=kappa,
                                    L_c0=f, x0=x0)
    assert rel(vals, ra) < TOL
    assert rel(b.get_local(), rb) < TOL
    # the vector-only variant
    asm = dfa.SystemAssembler(a, L, bc, scatter=scatter)
    b2 = dfa.Vector()
    asm.assemble(b=b2, x0=x0t)
    assert rel(b2.get_local(), rb) < TOL


def test_vp2_system_x0_vs_oracle(oracle):
    mesh, V, mu, lmbda, f3 = workloads.vp2_elasticity_3d(3, seed=14)
    Q = V.coefficient_space()
    a = dfa.elasticity_form(V, workloads.to_function(Q, mu), workloads.to_function(Q, lmbda))
    L = dfa.body_force_form(V, f3)
    rng = np.random.default_rng(15)
    bc = dfa.DirichletBC(V, 0.3)
    x0 = rng.uniform(-1, 1, V.dim())
    for scatter in (dfa.SCATTER_BLOCKED, dfa.SCATTER_ATOMIC):
        A, b = dfa.assemble_system(a, L, bc, x0=torch.from_numpy(x0).cuda(), scatter=scatter)
        rp, col, vals = A.data()
        mask = np.zeros(V.dim(), np.uint8)
        mask[bc.dofs] = 1
        ra, rb = oracle.assemble_system(oracle.VP2_ELASTICITY, oracle.VP2_BODY_FORCE,
                                        mesh.coordinates(), mesh.cells(), V.cell_dofs(), rp, col,
                                        mask, np.where(mask > 0, 0.3, 0.0),
                                        coeff_dofs=Q.cell_dofs(), a_c0=mu, a_c1=lmbda, f3=f3, x0=x0)
        assert rel(vals, ra) < TOL
        assert rel(b.get_local(), rb) < TOL


def test_p1_system_facets_x0_vs_oracle(oracle):
    # x0 with exterior-facet terms: the cell and facet passes share the shifted values
    mesh, V, kappa, f = workloads.p1_poisson_3d(4, seed=16)
    rng = np.random.default_rng(17)
    g, x0, gb = (rng.uniform(-1, 1, V.dim()) for _ in range(3))
    x = mesh.coordinates()
    cd = V.cell_dofs()
    low = np.nonzero(oracle.boundary_vertices(x) & (x[:, 0] < 0.5))[0]
    bc = dfa.DirichletBC(V, g[low], dofs=low)
    inside = lambda y: y[:, 0] > 0.3
    a = dfa.poisson_form(V, dfa.Function(V, kappa)) + dfa.facet_mass_form(V, 1.3, marker=inside)
    L = dfa.load_form(V, dfa.Function(V, f)) + dfa.facet_load_form(V, dfa.Function(V, gb), marker=inside)
    A, b = dfa.assemble_system(a, L, bc, x0=torch.from_numpy(x0).cuda())
    bcm = np.zeros(V.dim(), np.uint8)
    bcm[low] = 1
    rp, col = oracle.sparsity(cd, V.dim())
    Aref, bref = oracle.assemble_system_facets(
        oracle.P1_POISSON, oracle.P1_LOAD, oracle.P1_FACET_MASS, oracle.P1_FACET_LOAD, x,
        mesh.cells(), cd, oracle.exterior_facets(mesh.cells()), rp, col, bcm,
        np.where(bcm > 0, g, 0.0), coeff_dofs=cd, a_c0=kappa, L_c0=f, alpha=1.3, g=gb,
        facet_mask=dfa.mark_exterior_facets(V, inside), x0=x0)
    assert rel(A.values.cpu().numpy(), Aref) < TOL
    assert rel(b.get_local(), bref) < TOL


def test_x0_c_abi_entry_points(oracle):
    # dfa_assemble_system_x0 (one call) and dfa_dirichlet_x0 + dfa_assemble_system agree
    mesh, V, kappa, f = workloads.p1_poisson_3d(4, seed=18)
    rng = np.random.default_rng(19)
    eng = dfa.engine_for(V)
    lib = _lib.load()
    bc = dfa.DirichletBC(V, 0.0)
    g = torch.from_numpy(np.where(bc.mask.cpu().numpy() > 0, rng.uniform(-1, 1, V.dim()), 0.0)).cuda()
    x0 = torch.from_numpy(rng.uniform(-1, 1, V.dim())).cuda()
    K, F = torch.from_numpy(kappa).cuda(), torch.from_numpy(f).cuda()
    xd = mesh.coordinates_device()
    outs = []
    for one_call in (True, False):
        A = torch.zeros(eng.nnz, dtype=torch.float64, device="cuda")
        b = torch.zeros(V.dim(), dtype=torch.float64, device="cuda")
        if one_call:
            _lib.check(lib.dfa_assemble_system_x0(
                eng.handle, _lib.FORM_P1_POISSON, _lib.FORM_P1_LOAD, _lib.ptr(xd), _lib.ptr(K),
                None, _lib.ptr(F), None, _lib.ptr(bc.mask), _lib.ptr(g), _lib.ptr(x0), _lib.ptr(A),
                _lib.ptr(b), _lib.SCATTER_AUTO, 0, _lib.stream()))
        else:
            sh = torch.empty_like(g)
            _lib.check(lib.dfa_dirichlet_x0(V.dim(), _lib.ptr(bc.mask), _lib.ptr(g), _lib.ptr(x0),
                                            _lib.ptr(sh), _lib.stream()))
            m = bc.mask.bool()
            assert torch.equal(sh[m], (x0 - g)[m]) and torch.all(sh[~m] == 0)
            _lib.check(lib.dfa_assemble_system(
                eng.handle, _lib.FORM_P1_POISSON, _lib.FORM_P1_LOAD, _lib.ptr(xd), _lib.ptr(K),
                None, _lib.ptr(F), None, _lib.ptr(bc.mask), _lib.ptr(sh), _lib.ptr(A), _lib.ptr(b),
                _lib.SCATTER_AUTO, 0, _lib.stream()))
        outs.append((A, b))
    torch.cuda.synchronize()
    # same values; the LDS accumulation order of the blocked kernel may differ between passes
    for u, v in zip(outs[0], outs[1]):
        assert (torch.linalg.norm(u - v) / torch.linalg.norm(v)).item() < 1e-14
    # NULL x0 is refused by the one-call variant
    assert lib.dfa_assemble_system_x0(eng.handle, _lib.FORM_P1_POISSON, _lib.FORM_P1_LOAD,
                                      _lib.ptr(xd), _lib.ptr(K), None, _lib.ptr(F), None, None,
                                      None, None, _lib.ptr(outs[0][0]), None, 0, 0,
                                      _lib.stream()) != 0


def test_host_add_local_blocks(oracle):
    # GenericMatrix::add_local / GenericVector::add_local from host blocks: the cell loop of
    # Assembler::assemble_cells (Assembler.cpp:105-194) on the host, one add_local per cell,
    # reproduces the device assembly; an entry outside the pattern is refused
    mesh, V, kappa, f = workloads.p1_poisson_3d(2, seed=20)
    A = dfa.Matrix()
    A.init(dfa.engine_for(V))
    assert A.local_range(0) == (0, V.dim())
    b = dfa.Vector(V.dim())
    x = mesh.coordinates()
    for c, dofs in zip(mesh.cells(), V.cell_dofs()):
        Ke = oracle.p1_stiffness(x[c], kappa[dofs])
        be = oracle.p1_load(x[c], f[dofs])
        A.add_local(Ke, dofs, dofs)
        b.add_local(be, dofs)
    ref = dfa.assemble(dfa.poisson_form(V, workloads.to_function(V, kappa)))
    bref = dfa.assemble(dfa.load_form(V, workloads.to_function(V, f)))
    assert rel(A.values.cpu().numpy(), ref.values.cpu().numpy()) < TOL
    assert rel(b.get_local(), bref.get_local()) < TOL
    before = A.values.clone()
    with pytest.raises(RuntimeError, match="sparsity pattern"):
        A.add_local(np.ones((1, 1)), [0], [V.dim() - 1])
    assert torch.equal(A.values, before)


def test_pass_kernel_names():
    mesh, V, kappa, f = workloads.p1_poisson_3d(4, seed=21)
    eng = dfa.engine_for(V)
    P, Lf, M = _lib.FORM_P1_POISSON, _lib.FORM_P1_LOAD, _lib.FORM_P1_MASS
    assert eng.pass_kernel(P, -1).startswith("p1_chain_kernel<WANT_B=0")
    # Dirichlet conditions follow the edge-chain assembly (dirichlet_rows_kernel)
    assert eng.pass_kernel(P, Lf, has_bc=True) == "p1_chain_kernel<WANT_B=1,T=512,RR=1> + dirichlet_rows_kernel"
    assert eng.pass_kernel(M, -1).startswith("p1_blocked_kernel<D=3,AFORM=2,WANT_B=0,BC=0")
    assert eng.pass_kernel(P, Lf, scatter=dfa.SCATTER_ATOMIC) == "p1_atomic_kernel"
    assert eng.pass_kernel(P, Lf, scatter=dfa.SCATTER_COLORED) == "p1_colored_kernel"


def test_boundary_dofs_vector_p2():
    # DirichletBC 'on_boundary' on vector P2: every component of the vertex and edge nodes that
    # lie on the cube's faces (edge midpoints included), with UFC and permuted numbering
    for permuted in (False, True):
        mesh, V, mu, lmbda, f3 = _vp2_case(3, permuted, seed=22)
        x = mesh.coordinates()
        ce, ne = mesh.cell_edges()
        loc = np.array([[2, 3], [1, 3], [1, 2], [0, 3], [0, 2], [0, 1]])
        c = mesh.cells()
        pts = np.concatenate([x[c], 0.5 * (x[c[:, loc[:, 0]]] + x[c[:, loc[:, 1]]])], axis=1)
        on = np.any((np.abs(pts) < 1e-14) | (np.abs(pts - 1) < 1e-14), axis=2)   # [nc, 10]
        cd = V.cell_dofs()
        expect = np.unique(cd[np.concatenate([on] * 3, axis=1)])
        np.testing.assert_array_equal(dfa.fem.boundary_facet_nodes(V), expect)


@pytest.mark.parametrize("permuted", [False, True])
@pytest.mark.parametrize("n", [3, 9])
def test_p1_dirichlet_after_chains(oracle, monkeypatch, n, permuted):
    # SystemAssembler with Dirichlet conditions on the boundary and on scattered interior dofs:
    # the edge-chain kernel assembles without them and dirichlet_rows_kernel applies what the
    # per-cell elimination (SystemAssembler.cpp:1051-1075) sums up to. Checked against the
    # oracle's per-cell loop and against the per-cell elimination inside the blocked kernel
    # (DFA_BC_INLINE=1), for assemble(A, b), assemble(A) and the incremental assemble(A, b, x0).
    print(f"case dirichlet after chains n={n} permuted={permuted}", flush=True)
    mesh, V, kappa, f = _p1_case(n, permuted, seed=60 + n)
    rng = np.random.default_rng(61 + n)
    dofs = np.union1d(dfa.DirichletBC(V, 0.0).dofs, rng.choice(V.dim(), V.dim() // 7, replace=False))
    g = rng.uniform(-1, 1, V.dim())
    bc = dfa.DirichletBC(V, g[dofs], dofs)
    eng = dfa.engine_for(V)
    assert eng.pass_kernel(_lib.FORM_P1_POISSON, _lib.FORM_P1_LOAD, has_bc=True).endswith(
        "+ dirichlet_rows_kernel")
    a = dfa.poisson_form(V, workloads.to_function(V, kappa))
    L = dfa.load_form(V, workloads.to_function(V, f))
    asm = dfa.SystemAssembler(a, L, bc)
    A, b = dfa.Matrix(), dfa.Vector()
    asm.assemble(A, b)
    Aonly = dfa.Matrix()
    asm.assemble(Aonly)
    x0 = torch.from_numpy(rng.uniform(-1, 1, V.dim())).cuda()
    Ax, bx = dfa.Matrix(), dfa.Vector()
    asm.assemble(Ax, bx, x0)
    torch.cuda.synchronize()
    rp, col, vals = A.data()
    mask = np.zeros(V.dim(), np.uint8)
    mask[dofs] = 1
    gv = np.zeros(V.dim())
    gv[dofs] = g[dofs]
    ra, rb = oracle.assemble_system(oracle.P1_POISSON, oracle.P1_LOAD, mesh.coordinates(), mesh.cells(),
                                    V.cell_dofs(), rp, col, mask, gv, coeff_dofs=V.cell_dofs(),
                                    a_c0=kappa, L_c0=f)
    assert rel(vals, ra) < TOL
    assert rel(b.get_local(), rb) < TOL
    assert rel(Aonly.data()[2], ra) < TOL
    gx = np.zeros(V.dim())
    gx[dofs] = x0.cpu().numpy()[dofs] - g[dofs]
    _, rbx = oracle.assemble_system(oracle.P1_POISSON, oracle.P1_LOAD, mesh.coordinates(), mesh.cells(),
                                    V.cell_dofs(), rp, col, mask, gx, coeff_dofs=V.cell_dofs(),
                                    a_c0=kappa, L_c0=f)
    assert rel(Ax.data()[2], ra) < TOL
    assert rel(bx.get_local(), rbx) < TOL
    monkeypatch.setenv("DFA_BC_INLINE", "1")
    assert eng.pass_kernel(_lib.FORM_P1_POISSON, _lib.FORM_P1_LOAD, has_bc=True).startswith(
        "p1_blocked_kernel<D=3,AFORM=0,WANT_B=1,BC=1")
    A2, b2 = dfa.Matrix(), dfa.Vector()
    asm.assemble(A2, b2)
    torch.cuda.synchronize()
    assert rel(A2.data()[2], vals) < TOL
    assert rel(b2.get_local(), b.get_local()) < TOL


@pytest.mark.parametrize("n,permuted", [(3, False), (3, True), (6, False)])
def test_vp2_dirichlet_after_assembly(oracle, monkeypatch, n, permuted):
    # Vector P2 with Dirichlet conditions on the boundary and on scattered interior dofs: the
    # blocked kernel assembles without them and dirichlet_node_rows_kernel (then the free rows pull their
    # Dirichlet columns over the node graph) applies what the per-cell elimination
    # (SystemAssembler.cpp:1051-1075) sums up to. Checked against the oracle's per-cell loop,
    # the incremental form with x0, dirichlet_rows_kernel (DFA_BC_ROWS=1: pushed by the
    # Dirichlet rows, binary search for the transposed entry) and the per-cell elimination
    # inside the blocked kernel (DFA_BC_INLINE=1).
    print(f"case vp2 dirichlet after assembly n={n} permuted={permuted}", flush=True)
    mesh, V, mu, lmbda, f3 = _vp2_case(n, permuted, seed=70 + n)
    rng = np.random.default_rng(71 + n)
    dofs = np.union1d(dfa.DirichletBC(V, 0.0).dofs, rng.choice(V.dim(), V.dim() // 9, replace=False))
    g = rng.uniform(-1, 1, V.dim())
    bc = dfa.DirichletBC(V, g[dofs], dofs)
    eng = dfa.engine_for(V)
    kname = eng.pass_kernel(_lib.FORM_VP2_ELASTICITY, _lib.FORM_VP2_BODY_FORCE, has_bc=True)
    assert kname == "vp2_blocked_kernel<WANT_A=1,WANT_B=1,BC=0> + dirichlet_node_rows_kernel", kname
    Q = V.coefficient_space()
    a = dfa.elasticity_form(V, workloads.to_function(Q, mu), workloads.to_function(Q, lmbda))
    L = dfa.body_force_form(V, f3)
    asm = dfa.SystemAssembler(a, L, bc)
    A, b = dfa.Matrix(), dfa.Vector()
    asm.assemble(A, b)
    Aonly = dfa.Matrix()
    asm.assemble(Aonly)
    x0 = torch.from_numpy(rng.uniform(-1, 1, V.dim())).cuda()
    Ax, bx = dfa.Matrix(), dfa.Vector()
    asm.assemble(Ax, bx, x0)
    torch.cuda.synchronize()
    rp, col, vals = A.data()
    mask = np.zeros(V.dim(), np.uint8)
    mask[dofs] = 1
    kd = Q.cell_dofs()

    def ref(gvals):
        return oracle.assemble_system(oracle.VP2_ELASTICITY, oracle.VP2_BODY_FORCE,
                                      mesh.coordinates(), mesh.cells(), V.cell_dofs(), rp, col, mask,
                                      gvals, coeff_dofs=kd, a_c0=mu, a_c1=lmbda, f3=f3)

    gv = np.zeros(V.dim())
    gv[dofs] = g[dofs]
    ra, rb = ref(gv)
    assert rel(vals, ra) < TOL
    assert rel(b.get_local(), rb) < TOL
    assert rel(Aonly.data()[2], ra) < TOL
    gx = np.zeros(V.dim())
    gx[dofs] = x0.cpu().numpy()[dofs] - g[dofs]
    _, rbx = ref(gx)
    assert rel(Ax.data()[2], ra) < TOL
    assert rel(bx.get_local(), rbx) < TOL
    for env, expect in [("DFA_BC_ROWS", "vp2_blocked_kernel<WANT_A=1,WANT_B=1,BC=0> + dirichlet_rows_kernel"),
                        ("DFA_BC_INLINE", "vp2_blocked_kernel<WANT_A=1,WANT_B=1,BC=1>")]:
        monkeypatch.setenv(env, "1")
        assert eng.pass_kernel(_lib.FORM_VP2_ELASTICITY, _lib.FORM_VP2_BODY_FORCE, has_bc=True) == expect
        A2, b2 = dfa.Matrix(), dfa.Vector()
        asm.assemble(A2, b2)
        torch.cuda.synchronize()
        assert rel(A2.data()[2], vals) < TOL
        assert rel(b2.get_local(), b.get_local()) < TOL
        monkeypatch.delenv(env)


@pytest.mark.parametrize("element", ["p1", "vp2"])
def test_homogeneous_dirichlet_after_assembly(oracle, element):
    # g = 0 on the boundary: the post-passes zero the eliminated entries of the free rows
    # without reading them (nothing is lifted into b); A and b against the oracle's per-cell
    # elimination (SystemAssembler.cpp:1051-1075), for assemble(A, b) and assemble(A).
    print(f"case homogeneous dirichlet after assembly {element}", flush=True)
    if element == "p1":
        mesh, V, kappa, f = _p1_case(7, False, seed=80)
        a = dfa.poisson_form(V, workloads.to_function(V, kappa))
        L = dfa.load_form(V, workloads.to_function(V, f))
        kw = dict(coeff_dofs=V.cell_dofs(), a_c0=kappa, L_c0=f)
        forms = (oracle.P1_POISSON, oracle.P1_LOAD)
    else:
        mesh, V, mu, lmbda, f3 = _vp2_case(4, False, seed=81)
        Q = V.coefficient_space()
        a = dfa.elasticity_form(V, workloads.to_function(Q, mu), workloads.to_function(Q, lmbda))
        L = dfa.body_force_form(V, f3)
        kw = dict(coeff_dofs=Q.cell_dofs(), a_c0=mu, a_c1=lmbda, f3=f3)
        forms = (oracle.VP2_ELASTICITY, oracle.VP2_BODY_FORCE)
    bc = dfa.DirichletBC(V, 0.0)
    assert dfa.engine_for(V).pass_kernel(*(_lib.FORM_P1_POISSON, _lib.FORM_P1_LOAD) if element == "p1" else
                                         (_lib.FORM_VP2_ELASTICITY, _lib.FORM_VP2_BODY_FORCE),
                                         has_bc=True).split(" + ")[1].startswith("dirichlet_")
    asm = dfa.SystemAssembler(a, L, bc)
    A, b = dfa.Matrix(), dfa.Vector()
    asm.assemble(A, b)
    Aonly = dfa.Matrix()
    asm.assemble(Aonly)
    torch.cuda.synchronize()
    rp, col, vals = A.data()
    mask = np.zeros(V.dim(), np.uint8)
    mask[bc.dofs] = 1
    ra, rb = oracle.assemble_system(*forms, mesh.coordinates(), mesh.cells(), V.cell_dofs(), rp, col, mask,
                                    np.zeros(V.dim()), **kw)
    assert rel(vals, ra) < TOL
    assert rel(b.get_local(), rb) < TOL
    assert rel(Aonly.data()[2], ra) < TOL


@pytest.mark.parametrize("element", ["p1", "vp2"])
def test_keep_diagonal_dirichlet_after_assembly(oracle, element):
    # SystemAssembler with keep_diagonal (AssemblerBase.h:46-67): vector P2 then takes the
    # general CSR build (no node graph), so its conditions go through dirichlet_rows_kernel
    # with a binary search for each transposed entry; P1 keeps the transpose map. Random
    # values on the boundary and on scattered interior dofs, against the oracle.
    print(f"case keep_diagonal dirichlet after assembly {element}", flush=True)
    rng = np.random.default_rng(90)
    if element == "p1":
        mesh, V, kappa, f = _p1_case(6, True, seed=91)
        a = dfa.poisson_form(V, workloads.to_function(V, kappa))
        L = dfa.load_form(V, workloads.to_function(V, f))
        kw = dict(coeff_dofs=V.cell_dofs(), a_c0=kappa, L_c0=f)
        forms = (oracle.P1_POISSON, oracle.P1_LOAD)
        kinds = (_lib.FORM_P1_POISSON, _lib.FORM_P1_LOAD)
    else:
        mesh, V, mu, lmbda, f3 = _vp2_case(3, False, seed=92)
        Q = V.coefficient_space()
        a = dfa.elasticity_form(V, workloads.to_function(Q, mu), workloads.to_function(Q, lmbda))
        L = dfa.body_force_form(V, f3)
        kw = dict(coeff_dofs=Q.cell_dofs(), a_c0=mu, a_c1=lmbda, f3=f3)
        forms = (oracle.VP2_ELASTICITY, oracle.VP2_BODY_FORCE)
        kinds = (_lib.FORM_VP2_ELASTICITY, _lib.FORM_VP2_BODY_FORCE)
    dofs = np.union1d(dfa.DirichletBC(V, 0.0).dofs, rng.choice(V.dim(), V.dim() // 8, replace=False))
    g = rng.uniform(-1, 1, V.dim())
    bc = dfa.DirichletBC(V, g[dofs], dofs)
    kname = dfa.engine_for(V, keep_diagonal=True).pass_kernel(*kinds, has_bc=True)
    assert kname.endswith("+ dirichlet_rows_kernel"), kname
    asm = dfa.SystemAssembler(a, L, bc)
    asm.keep_diagonal = True
    A, b = dfa.Matrix(), dfa.Vector()
    asm.assemble(A, b)
    torch.cuda.synchronize()
    rp, col, vals = A.data()
    mask = np.zeros(V.dim(), np.uint8)
    mask[dofs] = 1
    gv = np.zeros(V.dim())
    gv[dofs] = g[dofs]
    ra, rb = oracle.assemble_system(*forms, mesh.coordinates(), mesh.cells(), V.cell_dofs(), rp, col, mask, gv,
                                    **kw)
    assert rel(vals, ra) < TOL
    assert rel(b.get_local(), rb) < TOL
