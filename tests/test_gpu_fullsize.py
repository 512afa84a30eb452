"""Full-size checks of the benchmark workloads, through size-independent properties.

  P1 128^3 (12,582,912 tets): the fused A+b kernel the benchmark times, against the global-
      atomic scatter (a different kernel and summation order); b = M f with the P1 mass
      matrix of the per-cell blocked kernel and the device SpMV (f is P1, so f*v*dx is M f);
      K 1 = 0; symmetry.
  vector P2 64^3 (1,572,864 tets, nnz 547,739,145): the fused A+b kernel the benchmark
      times, against the atomic scatter; the six rigid-body modes in the null space; symmetry;
      sum of the z-load = -volume, x and y loads zero.
  Batch, 256 UnitCubeMesh members n ~ U{8..48} (seed 5, the benchmark's batch): every member
      of the graph-captured launch against its own standalone assembly, K 1 = 0 (P1) or the
      rigid modes (vector P2) on every member, and the oracle on the smallest P1 and P2
      members.
"""
import numpy as np
import pytest
import torch

import dolfin_amd as dfa
from dolfin_amd import _lib, workloads
from dolfin_amd.batch import BatchAssembler

pytestmark = pytest.mark.gpu

TOL = 1e-12


def trel(a, b):
    return (torch.linalg.norm(a - b) / torch.linalg.norm(b)).item()


def _symmetric(A, n, reps=3, seed=0):
    g = torch.Generator(device="cuda").manual_seed(seed)
    worst = 0.0
    for _ in range(reps):
        u = torch.rand(n, dtype=torch.float64, device="cuda", generator=g) - 0.5
        v = torch.rand(n, dtype=torch.float64, device="cuda", generator=g) - 0.5
        a = torch.dot(v, A.mult(u)).item()
        b = torch.dot(u, A.mult(v)).item()
        worst = max(worst, abs(a - b) / (abs(a) + abs(b) + 1e-300))
    return worst


def test_p1_128_fused_system():
    mesh, V, kappa, f = workloads.p1_poisson_3d(128)
    print("case p1 128 fused system", flush=True)
    eng = dfa.engine_for(V)
    assert eng.pass_kernel(_lib.FORM_P1_POISSON, _lib.FORM_P1_LOAD) == \
        "p1_chain_kernel<WANT_B=1,T=512,RR=1>"
    K, F = workloads.to_function(V, kappa), workloads.to_function(V, f)
    a, L = dfa.poisson_form(V, K), dfa.load_form(V, F)
    A, b = dfa.assemble_system(a, L)
    Aa = dfa.assemble(a, scatter=dfa.SCATTER_ATOMIC)
    ba = dfa.assemble(L, scatter=dfa.SCATTER_ATOMIC)
    assert trel(A.values, Aa.values) < TOL
    assert trel(b.values, ba.values) < TOL
    del Aa
    M = dfa.assemble(dfa.mass_form(V))
    assert eng.pass_kernel(_lib.FORM_P1_MASS, -1).startswith("p1_blocked_kernel")
    Mf = M.mult(F.vector())
    assert trel(b.values, Mf) < TOL
    # sum(b) = int f_h = 1^T M f
    s = b.values.sum().item()
    assert abs(s - torch.dot(M.mult(torch.ones_like(Mf)), F.vector()).item()) < 1e-12 * abs(s) + 1e-15
    one = torch.ones(V.dim(), dtype=torch.float64, device="cuda")
    K1 = A.mult(one)
    assert (K1.abs().max() / A.diagonal().abs().max()).item() < 1e-13
    assert _symmetric(A, V.dim()) < 1e-13


def test_vp2_64_fused_system():
    mesh, V, mu, lmbda, f3 = workloads.vp2_elasticity_3d(64)
    print("case vp2 64 fused system", flush=True)
    eng = dfa.engine_for(V)
    assert eng.pass_kernel(_lib.FORM_VP2_ELASTICITY, _lib.FORM_VP2_BODY_FORCE) == \
        "vp2_blocked_kernel<WANT_A=1,WANT_B=1,BC=0>"
    Q = V.coefficient_space()
    a = dfa.elasticity_form(V, workloads.to_function(Q, mu), workloads.to_function(Q, lmbda))
    L = dfa.body_force_form(V, f3)
    A, b = dfa.assemble_system(a, L)
    torch.cuda.synchronize()
    print("  assembled (fused)", flush=True)
    Aa = dfa.assemble(a, scatter=dfa.SCATTER_ATOMIC)
    ba = dfa.assemble(L, scatter=dfa.SCATTER_ATOMIC)
    assert trel(A.values, Aa.values) < TOL
    assert trel(b.values, ba.values) < TOL
    del Aa, ba
    torch.cuda.empty_cache()
    print("  blocked == atomic", flush=True)
    scale = A.values.abs().max().item()
    for r in workloads.rigid_body_modes(V):
        rt = torch.from_numpy(r).cuda()
        Kr = A.mult(rt)
        assert Kr.abs().max().item() < 1e-11 * scale * max(1.0, float(np.abs(r).max()))
    assert _symmetric(A, V.dim()) < 1e-13
    # body force (0, 0, -1): the z-components sum to -volume, the others to 0
    comp = np.zeros(V.dim(), np.int64)
    cd = V.cell_dofs()
    for i in range(3):
        comp[cd[:, 10 * i:10 * i + 10].ravel()] = i
    compt = torch.from_numpy(comp).cuda()
    bz = b.values[compt == 2].sum().item()
    bxy = b.values[compt != 2].abs().max().item()
    assert abs(bz + 1.0) < 1e-12 and bxy == 0.0


def _member_checks(a, A):
    """Standalone assembly of the member, and its null space."""
    ref = dfa.assemble(a)
    e = trel(A.values, ref.values)
    del ref
    V = a.V
    if a.kind == _lib.FORM_P1_POISSON:
        one = torch.ones(V.dim(), dtype=torch.float64, device="cuda")
        null = (A.mult(one).abs().max() / A.diagonal().abs().max()).item()
        return e, null < 1e-13
    scale = A.values.abs().max().item()
    ok = True
    for r in workloads.rigid_body_modes(V):
        Kr = A.mult(torch.from_numpy(r).cuda())
        ok &= Kr.abs().max().item() < 1e-11 * scale * max(1.0, float(np.abs(r).max()))
    return e, ok


@pytest.mark.timeout(900)
def test_batch_256_members(oracle):
    print("case batch 256 members seed 5", flush=True)
    forms = workloads.batch_members(256, seed=5, log=lambda m: print(m, flush=True))
    batch = BatchAssembler(forms)
    print(f"  batch created {batch.info}", flush=True)
    tensors = batch.assemble()
    torch.cuda.synchronize()
    bad = []
    for i, (a, A) in enumerate(zip(forms, tensors)):
        e, null_ok = _member_checks(a, A)
        if not (e < TOL and null_ok):
            bad.append((i, a.kind, a.V.dim(), e, null_ok))
        if (i + 1) % 32 == 0:
            print(f"  {i + 1}/256 members checked", flush=True)
    assert not bad, bad
    # the smallest P1 and vector-P2 members against the oracle
    for kind in (_lib.FORM_P1_POISSON, _lib.FORM_VP2_ELASTICITY):
        i = min((j for j, a in enumerate(forms) if a.kind == kind), key=lambda j: forms[j].V.dim())
        a, A = forms[i], tensors[i]
        mesh = a.V.mesh()
        rp, col, vals = A.data()
        cv = [c.cpu().numpy() for c in a.coefficient_vectors()]
        if kind == _lib.FORM_P1_POISSON:
            ref = oracle.assemble_matrix(oracle.P1_POISSON, mesh.coordinates(), mesh.cells(),
                                         a.V.cell_dofs(), rp, col, c0=cv[0])
        else:
            ref = oracle.assemble_matrix(oracle.VP2_ELASTICITY, mesh.coordinates(), mesh.cells(),
                                         a.V.cell_dofs(), rp, col,
                                         coeff_dofs=a.V.coefficient_space().cell_dofs(),
                                         c0=cv[0], c1=cv[1])
        assert np.linalg.norm(vals - ref) / np.linalg.norm(ref) < TOL, (i, kind)


def test_batch_block_body_member(oracle):
    # a P1 member whose row blocks hold more than 512 edge chains runs the per-cell block body
    # inside the batched P1 kernel (the chain body takes one chain per thread)
    from dolfin_amd import autotune
    autotune.set_plan_defaults(320, 0, 0)
    try:
        mesh, V, kappa, f = workloads.p1_poisson_3d(18, seed=23)  # 549 chains in the largest block
        eng = dfa.engine_for(V)
    finally:
        autotune.set_plan_defaults(0, 0, 0)
    big = eng.pass_kernel(_lib.FORM_P1_POISSON, -1)
    assert big == "p1_chain_kernel<WANT_B=0,T=512,RR=2>", big  # max_chains in (512, 1024]
    forms = [dfa.poisson_form(V, workloads.to_function(V, kappa))]
    m2, V2, k2, f2 = workloads.p1_poisson_3d(9, seed=24)
    forms.append(dfa.poisson_form(V2, workloads.to_function(V2, k2)))
    m3, V3, mu, lmbda, f3 = workloads.vp2_elasticity_3d(3, seed=25)
    Q = V3.coefficient_space()
    forms.append(dfa.elasticity_form(V3, workloads.to_function(Q, mu), workloads.to_function(Q, lmbda)))
    batch = BatchAssembler(forms)
    tensors = batch.assemble()
    torch.cuda.synchronize()
    for a, A in zip(forms, tensors):
        e, null_ok = _member_checks(a, A)
        assert e < TOL and null_ok
    rp, col, vals = tensors[0].data()
    ref = oracle.assemble_matrix(oracle.P1_POISSON, mesh.coordinates(), mesh.cells(),
                                 V.cell_dofs(), rp, col, c0=kappa)
    assert np.linalg.norm(vals - ref) / np.linalg.norm(ref) < TOL
