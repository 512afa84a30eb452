"""Batched assembly (one persistent, graph-captured launch over many meshes) against the
individual assemblies and the oracle; graph replays with coefficient updates between them."""
import numpy as np
import pytest
import torch

import dolfin_amd as dfa
from dolfin_amd import workloads
from dolfin_amd.batch import BatchAssembler

pytestmark = pytest.mark.gpu


def rel(a, b):
    return (torch.linalg.norm(a - b) / torch.linalg.norm(b)).item()


def _members(seed=0):
    rng = np.random.default_rng(seed)
    forms = []
    for n in (4, 7, 10):
        mesh, V, kappa, f = workloads.p1_poisson_3d(n, seed=int(rng.integers(1 << 30)))
        forms.append(dfa.poisson_form(V, workloads.to_function(V, kappa)))
    for n in (2, 3, 5):
        mesh, V, mu, lmbda, f3 = workloads.vp2_elasticity_3d(n, seed=int(rng.integers(1 << 30)))
        Q = V.coefficient_space()
        forms.append(dfa.elasticity_form(V, workloads.to_function(Q, mu), workloads.to_function(Q, lmbda)))
    return forms


def test_batch_matches_individual_and_oracle(oracle):
    forms = _members()
    batch = BatchAssembler(forms)
    print(f"case batch info={batch.info}", flush=True)
    tensors = batch.assemble()
    torch.cuda.synchronize()
    for a, A in zip(forms, tensors):
        ref = dfa.assemble(a, scatter=dfa.SCATTER_ATOMIC)
        assert rel(A.values, ref.values) < 1e-12
    a0 = forms[0]
    mesh = a0.V.mesh()
    rp, col, vals = tensors[0].data()
    kappa = a0.coefficients[0].vector().cpu().numpy()
    ref = oracle.assemble_matrix(oracle.P1_POISSON, mesh.coordinates(), mesh.cells(), a0.V.cell_dofs(),
                                 rp, col, c0=kappa)
    assert np.linalg.norm(vals - ref) / np.linalg.norm(ref) < 1e-12


def test_batch_graph_replays_follow_coefficients():
    forms = _members(1)
    batch = BatchAssembler(forms, grid_per_cu=1)
    first = [A.values.clone() for A in batch.assemble()]
    for rep in range(3):
        # eager work between replays: scale every coefficient, then replay the graph
        s = 1.5 + rep
        for a in forms:
            for c in a.coefficients:
                c.vector().mul_(s)
        torch.cuda.synchronize()
        tensors = batch.assemble()
        torch.cuda.synchronize()
        for a, A in zip(forms, tensors):
            ref = dfa.assemble(a, scatter=dfa.SCATTER_ATOMIC)
            assert rel(A.values, ref.values) < 1e-12, rep
    # the forms are linear in their coefficients: the last replay is the first times prod(s)
    total = 1.5 * 2.5 * 3.5
    for A, v in zip(batch.tensors, first):
        assert rel(A.values, total * v) < 1e-12
