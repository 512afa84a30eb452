"""CPU tests of the host-side structure phase and host logic (no GPU calls):
CSR structure bit-exact against the oracle's SparsityPattern restatement, the blocked-scatter
partition walked index by index (tests/plan_check.cpp), product meshes/dofmaps against the
oracle's generators, the C-ABI library exports."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT, host_csr, p

import dolfin_amd as dfa
from dolfin_amd import workloads


def test_meshes_match_oracle(oracle):
    m = dfa.UnitCubeMesh(3, 4, 2)
    x, c = oracle.box_mesh(3, 4, 2)
    np.testing.assert_array_equal(m.coordinates(), x)
    np.testing.assert_array_equal(m.cells(), c)
    m = dfa.UnitSquareMesh(5, 3)
    x, c = oracle.rectangle_mesh_right(5, 3)
    np.testing.assert_array_equal(m.coordinates(), x)
    np.testing.assert_array_equal(m.cells(), c)
    m = dfa.UnitIntervalMesh(7)
    x, c = oracle.interval_mesh(7)
    np.testing.assert_array_equal(m.coordinates(), x)
    np.testing.assert_array_equal(m.cells(), c)


def test_vp2_dofmap_matches_oracle(oracle):
    m = dfa.UnitCubeMesh(2, 3, 2)
    V = dfa.VectorFunctionSpace(m, "P", 2)
    ce, ne = oracle.tet_edges(m.cells())
    np.testing.assert_array_equal(m.cell_edges()[0], ce)
    cd = oracle.vp2_dofmap(m.cells(), ce, m.num_vertices(), ne)
    np.testing.assert_array_equal(V.cell_dofs(), cd)
    assert V.dim() == 3 * (m.num_vertices() + ne)


@pytest.mark.parametrize("case", ["square", "cube", "interval", "cube_perm", "vp2", "vp2_block"])
def test_csr_structure_bit_exact(oracle, plan_lib, case):
    rng = np.random.default_rng(1)
    if case == "square":
        V = dfa.FunctionSpace(dfa.UnitSquareMesh(64, 64), "P", 1)
    elif case == "cube":
        V = dfa.FunctionSpace(dfa.UnitCubeMesh(5, 4, 3), "P", 1)
    elif case == "interval":
        V = dfa.FunctionSpace(dfa.UnitIntervalMesh(33), "P", 1)
    elif case == "cube_perm":
        m = dfa.UnitCubeMesh(4, 4, 4)
        V = dfa.FunctionSpace(m, "P", 1, node_permutation=rng.permutation(m.num_vertices()))
    elif case == "vp2":
        V = dfa.VectorFunctionSpace(dfa.UnitCubeMesh(3, 2, 2), "P", 2)
    else:
        m = dfa.UnitCubeMesh(2, 2, 2)
        nn = m.num_vertices() + m.cell_edges()[1]
        V = dfa.VectorFunctionSpace(m, "P", 2, node_permutation=rng.permutation(nn))
    rp, col = host_csr(plan_lib, V.cell_dofs(), V.dim())
    orp, ocol = oracle.sparsity(V.cell_dofs(), V.dim())
    np.testing.assert_array_equal(rp, orp)
    np.testing.assert_array_equal(col, ocol)
    if case == "square":
        assert col.shape[0] == 29057


def test_csr_keep_diagonal(oracle, plan_lib):
    V = dfa.FunctionSpace(dfa.UnitSquareMesh(3, 3), "P", 1)
    rp, col = host_csr(plan_lib, V.cell_dofs(), V.dim() + 3, keep_diagonal=True)
    orp, ocol = oracle.sparsity(V.cell_dofs(), V.dim() + 3, keep_diagonal=True)
    np.testing.assert_array_equal(rp, orp)
    np.testing.assert_array_equal(col, ocol)


def _plan_check(plan_lib, V, target_rows):
    mesh = V.mesh()
    cells = np.ascontiguousarray(mesh.cells(), np.int32)
    cd = np.ascontiguousarray(V.cell_dofs(), np.int32)
    x = np.ascontiguousarray(mesh.coordinates())
    cap = cells.shape[0] * cells.shape[1] ** 2
    vals = np.zeros(cap)
    rp = np.zeros(V.dim() + 1, np.int32)
    stats = np.zeros(8, np.int32)
    rc = plan_lib.plan_check_p1(mesh.gdim(), mesh.num_vertices(), mesh.num_cells(), p(cells), p(cd),
                                p(cd), p(x), V.dim(), target_rows, 0, p(vals), p(rp), cap, p(stats))
    assert rc == 0, f"plan check failed with code {rc}"
    return rp, vals[: rp[-1]], stats


@pytest.mark.parametrize("case,target", [("square", 64), ("square", 512), ("cube", 32),
                                         ("cube", 256), ("cube_perm", 100), ("interval", 8),
                                         ("delaunay", 64), ("delaunay", 160)])
def test_blocked_plan_walk(oracle, plan_lib, case, target):
    rng = np.random.default_rng(2)
    if case == "square":
        mesh = dfa.UnitSquareMesh(32, 32)
    elif case == "interval":
        mesh = dfa.UnitIntervalMesh(50)
    elif case == "delaunay":
        # unstructured tetrahedra: rows longer than 16 entries, 8-bit ranks
        import scipy.spatial
        pts = np.vstack([rng.uniform(0, 1, (600, 3)),
                         np.array(np.meshgrid([0., 1.], [0., 1.], [0., 1.])).reshape(3, -1).T])
        mesh = dfa.Mesh(pts, np.sort(scipy.spatial.Delaunay(pts).simplices, axis=1))
    else:
        mesh = dfa.UnitCubeMesh(6, 5, 7)
        dfa.jitter_interior(mesh, 0.2 / 7, rng)
    perm = rng.permutation(mesh.num_vertices()) if case == "cube_perm" else None
    V = dfa.FunctionSpace(mesh, "P", 1, node_permutation=perm)
    rp, vals, stats = _plan_check(plan_lib, V, target)
    orp, ocol = oracle.sparsity(V.cell_dofs(), V.dim())
    ref = oracle.assemble_matrix(oracle.P1_MASS, mesh.coordinates(), mesh.cells(), V.cell_dofs(),
                                 orp, ocol)
    np.testing.assert_array_equal(rp, orp)
    assert np.linalg.norm(vals - ref) <= 1e-13 * np.linalg.norm(ref)
    assert stats[2] <= target


@pytest.mark.parametrize("n,target,perm", [(6, 32, False), (7, 160, True), (12, 160, False)])
def test_chain_plan_walk(oracle, plan_lib, n, target, perm):
    # edge chains of the tetrahedral kappa-Poisson kernel, walked on the CPU exactly as a lane
    # does it, must reproduce the oracle's stiffness matrix and load vector
    rng = np.random.default_rng(5)
    mesh = dfa.UnitCubeMesh(n, n + 1, n - 1)
    dfa.jitter_interior(mesh, 0.2 / n, rng)
    V = dfa.FunctionSpace(mesh, "P", 1,
                          node_permutation=rng.permutation(mesh.num_vertices()) if perm else None)
    kappa = rng.uniform(0.5, 2.0, V.dim())
    f = rng.uniform(-1.0, 1.0, V.dim())
    cells = np.ascontiguousarray(mesh.cells(), np.int32)
    cd = np.ascontiguousarray(V.cell_dofs(), np.int32)
    x = np.ascontiguousarray(mesh.coordinates())
    orp, ocol = oracle.sparsity(V.cell_dofs(), V.dim())
    vals = np.zeros(ocol.shape[0])
    b = np.zeros(V.dim())
    stats = np.zeros(8, np.int32)
    rc = plan_lib.plan_check_p1_chains(mesh.num_vertices(), mesh.num_cells(), p(cells), p(cd),
                                       p(x), p(kappa), p(f), V.dim(), target, p(vals), p(b),
                                       ocol.shape[0], p(stats))
    assert rc == 0, f"chain plan check failed with code {rc}"
    A_ref = oracle.assemble_matrix(oracle.P1_POISSON, mesh.coordinates(), mesh.cells(), cd, orp,
                                   ocol, coeff_dofs=cd, c0=kappa)
    b_ref = oracle.assemble_vector(oracle.P1_LOAD, mesh.coordinates(), mesh.cells(), cd, V.dim(),
                                   coeff_dofs=cd, c0=f)
    assert np.linalg.norm(vals - A_ref) <= 1e-12 * np.linalg.norm(A_ref)
    assert np.linalg.norm(b - b_ref) <= 1e-12 * np.linalg.norm(b_ref)
    # chains average well above 2 cells on a Kuhn mesh
    assert stats[3] / stats[2] > 2.5


def test_blocked_plan_3d_redundancy(plan_lib):
    # the work partition lists each cell in every block owning one of its dofs; on a
    # structured 3D mesh with ~256-row blocks that is well under 2x
    mesh = dfa.UnitCubeMesh(16, 16, 16)
    V = dfa.FunctionSpace(mesh, "P", 1)
    _, _, stats = _plan_check(plan_lib, V, 256)
    assert stats[4] / mesh.num_cells() < 2.0


def test_xcd_block_mapping_is_bijective():
    # p1_blocked_kernel: blk = xcd*q + min(xcd, r) + j with xcd = i % 8, j = i // 8
    for nb in [1, 7, 8, 9, 63, 64, 65, 1000, 8385]:
        q, r = nb // 8, nb % 8
        i = np.arange(nb)
        xcd, j = i % 8, i // 8
        blk = xcd * q + np.minimum(xcd, r) + j
        assert np.array_equal(np.sort(blk), np.arange(nb))


@pytest.mark.parametrize("case", ["p1_cube", "fan", "empty_rows"])
def test_spmv_row_blocks(plan_lib, case):
    # row blocks of the streamed SpMV: cover every row once, in order; <= 256 rows and
    # <= 4096 nonzeros per block unless the block is one longer row; maximal (the next row
    # would break a limit)
    if case == "p1_cube":
        rp = host_csr(plan_lib, dfa.UnitCubeMesh(9, 7, 8).cells(), 720)[0]
    elif case == "fan":
        lens = np.r_[np.full(300, 4), 5001, np.full(700, 4), 9000, 4096, 4097, np.full(600, 15)]
        rp = np.r_[0, np.cumsum(lens)].astype(np.int32)
    else:
        lens = np.random.default_rng(2).integers(0, 3, 5000)
        rp = np.r_[0, np.cumsum(lens)].astype(np.int32)
    n = rp.shape[0] - 1
    nb = plan_lib.spmv_blocks_host(p(rp), n, 4096, 256, None)
    b = np.empty(nb + 1, np.int32)
    plan_lib.spmv_blocks_host(p(rp), n, 4096, 256, p(b))
    assert b[0] == 0 and b[-1] == n and np.all(np.diff(b) > 0)
    for r0, r1 in zip(b[:-1], b[1:]):
        nnz = rp[r1] - rp[r0]
        assert r1 - r0 <= 256
        assert nnz <= 4096 or r1 - r0 == 1
        if r1 < n and nnz <= 4096:
            assert r1 - r0 == 256 or rp[r1 + 1] - rp[r0] > 4096


def test_library_exports_header_symbols():
    from dolfin_amd import _lib
    hdr = open(os.path.join(ROOT, "include", "dolfin_amd.h")).read()
    declared = set(re.findall(r"^\s*(?:int|const char\*)\s+(dfa_\w+)\s*\(", hdr, re.M))
    assert declared == set(_lib.EXPORTS)
    L = _lib.load()
    for name in declared:
        assert hasattr(L, name), name
    assert L.dfa_version().decode().startswith("dolfin_amd")


def test_errors_without_gpu_calls():
    from dolfin_amd import _lib
    L = _lib.load()
    h = ctypes.c_void_p()
    cells = np.zeros((1, 4), np.int32)
    rc = L.dfa_assembler_create(ctypes.byref(h), 4, 1, 1, p(cells), 1, 5, p(cells), 1, p(cells), 1,
                                None, 0, None)
    assert rc != 0 and "gdim" in L.dfa_last_error().decode()
    rc = L.dfa_assembler_create(ctypes.byref(h), 3, 4, 1, p(cells), 7, 4, p(cells), 4, p(cells), 4,
                                None, 0, None)
    assert rc != 0 and "element" in L.dfa_last_error().decode()
    bad = np.array([[0, 1, 2, 9]], np.int32)
    rc = L.dfa_assembler_create(ctypes.byref(h), 3, 4, 1, p(bad), 1, 4, p(bad), 4, p(bad), 4,
                                None, 0, None)
    assert rc != 0 and "out of range" in L.dfa_last_error().decode()


def test_workload_sizes():
    mesh, V, kappa, f = workloads.p1_poisson_2d(64)
    assert mesh.num_cells() == 8192 and V.dim() == 4225
    # algorithmic bytes of the 2D parity config (12 B connectivity/cell)
    assert workloads.p1_algorithmic_bytes(8192, 3, 4225, 2, 29057, 4225) == 565292
    assert workloads.p1_algorithmic_bytes(12582912, 4, 2146689, 3, 31802497, 2146689) == 660237364


@pytest.mark.parametrize("n,block,vals", [(2, False, 2048), (3, False, 6144), (3, True, 4096),
                                          (4, False, 6144)])
def test_vp2_blocked_plan_walk(oracle, plan_lib, n, block, vals):
    rng = np.random.default_rng(n)
    mesh = dfa.UnitCubeMesh(n, n, n)
    dfa.jitter_interior(mesh, 0.15 / n, rng)
    perm = None
    if block:
        perm = rng.permutation(mesh.num_vertices() + mesh.cell_edges()[1])
    V = dfa.VectorFunctionSpace(mesh, "P", 2, node_permutation=perm)
    Q = V.coefficient_space()
    mu = rng.uniform(0.5, 1.5, Q.dim())
    la = rng.uniform(1, 10, Q.dim())
    cells = np.ascontiguousarray(mesh.cells(), np.int32)
    cd = np.ascontiguousarray(V.cell_dofs(), np.int32)
    kd = np.ascontiguousarray(Q.cell_dofs(), np.int32)
    x = np.ascontiguousarray(mesh.coordinates())
    cap = cells.shape[0] * 900
    out = np.zeros(cap)
    stats = np.zeros(8, np.int32)
    rc = plan_lib.plan_check_vp2(cells.shape[0], p(cells), p(cd), p(kd), p(x), V.dim(), p(mu), p(la),
                                 vals, p(out), cap, p(stats))
    assert rc == 0, f"vector P2 plan check failed with code {rc}"
    rp, col = oracle.sparsity(cd, V.dim())
    ref = oracle.assemble_matrix(oracle.VP2_ELASTICITY, x, cells, cd, rp, col, coeff_dofs=kd, c0=mu, c1=la)
    got = out[: col.shape[0]]
    assert np.linalg.norm(got - ref) <= 1e-13 * np.linalg.norm(ref)
    assert stats[3] <= vals


def test_cpp_adapter_example_is_built():
    # built by dolfin_amd.build (host C++ against include/dolfin_amd/HipAssembler.hpp)
    path = os.path.join(ROOT, "tests", "_build", "cpp_system_assembler")
    assert os.path.exists(path) and os.access(path, os.X_OK)
