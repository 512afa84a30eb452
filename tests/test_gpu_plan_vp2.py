"""The device planner of the vector-P2 blocked scatter against the host builder.

With DFA_PLAN_VERIFY=1 in the environment, creating a vector-P2 assembler builds the rank
table and the block plan both on the device (plan_vp2_device.hip) and on the host
(structure.cpp) and compares every array element by element; plan_builder() raises on any
difference and names the first differing array. The cases cover the recursive-bisection ties
of an unjittered grid (equal coordinates are ordered by node), a stretched box (the split axis
changes from level to level), both node-blocked dof layouts (component-major UFC numbering
and node-major after a node permutation or reverse Cuthill-McKee), the block sizes an
autotuner may choose, and the 64^3 benchmark mesh. Without the flag the device plan is the
one assembled with; the assembled values are checked here against the oracle as well.
"""
import numpy as np
import pytest
import torch

import dolfin_amd as dfa
from dolfin_amd import autotune, workloads

pytestmark = pytest.mark.gpu


def _space(n, seed, jitter=0.15, stretch=None, node_major=False, rcm=False):
    rng = np.random.default_rng(seed)
    mesh = dfa.UnitCubeMesh(n, n, n)
    if jitter:
        dfa.jitter_interior(mesh, jitter / n, rng)
    if stretch is not None:
        mesh.set_coordinates(mesh.coordinates() * np.asarray(stretch))
    perm = None
    if node_major:
        perm = rng.permutation(mesh.num_vertices() + mesh.cell_edges()[1]).astype(np.int32)
    if rcm:
        ce, ne = mesh.cell_edges()
        nodes = np.concatenate([mesh.cells(), mesh.num_vertices() + ce], axis=1)
        perm = dfa.node_reordering(nodes, mesh.num_vertices() + ne)
    return mesh, dfa.VectorFunctionSpace(mesh, "P", 2, node_permutation=perm)


CASES = {
    "n1": dict(n=1, seed=1),
    "n2": dict(n=2, seed=2),
    "n5": dict(n=5, seed=3),
    "n11": dict(n=11, seed=4),
    "n6_grid_ties": dict(n=6, seed=5, jitter=0.0),
    "n9_grid_ties": dict(n=9, seed=6, jitter=0.0),
    "n8_stretched": dict(n=8, seed=7, stretch=(4.0, 1.0, 0.25)),
    "n7_node_major": dict(n=7, seed=8, node_major=True),
    "n8_rcm": dict(n=8, seed=9, rcm=True),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_device_vp2_plan_equals_host_plan(monkeypatch, name):
    print(f"case device vp2 plan {name}", flush=True)
    monkeypatch.setenv("DFA_PLAN_VERIFY", "1")
    mesh, V = _space(**CASES[name])
    assert dfa.fem.Engine(V).plan_builder() == "device"


@pytest.mark.parametrize("vals", [1024, 2048, 6144, 16384])
def test_device_vp2_plan_block_sizes(monkeypatch, vals):
    print(f"case device vp2 plan block values {vals}", flush=True)
    monkeypatch.setenv("DFA_PLAN_VERIFY", "1")
    autotune.set_plan_defaults(0, 0, vals)
    try:
        mesh, V = _space(8, seed=20 + vals % 7)
        assert dfa.fem.Engine(V).plan_builder() == "device"
    finally:
        autotune.set_plan_defaults(0, 0, 0)


def test_host_plan_switch(monkeypatch):
    monkeypatch.setenv("DFA_HOST_PLAN", "1")
    mesh, V = _space(3, seed=30)
    assert dfa.fem.Engine(V).plan_builder() == "host"


@pytest.mark.parametrize("node_major", [False, True])
def test_device_vp2_plan_assembles_like_the_oracle(oracle, node_major):
    print(f"case device vp2 plan assembly node_major={node_major}", flush=True)
    mesh, V = _space(4, seed=31, node_major=node_major)
    eng = dfa.fem.Engine(V)
    assert eng.plan_builder() == "device"
    rng = np.random.default_rng(32)
    Q = V.coefficient_space()
    mu, lmbda = rng.uniform(0.5, 1.5, Q.dim()), rng.uniform(1.0, 10.0, Q.dim())
    f3 = np.array([0.3, -0.2, -1.0])
    a = dfa.elasticity_form(V, workloads.to_function(Q, mu), workloads.to_function(Q, lmbda))
    L = dfa.body_force_form(V, f3)
    for scatter in (dfa.SCATTER_BLOCKED, dfa.SCATTER_ATOMIC):  # the atomic scatter reads the ranks
        A = dfa.assemble(a, scatter=scatter)
        b = dfa.assemble(L, scatter=scatter)
        torch.cuda.synchronize()
        rp, col, vals = A.data()
        kd = Q.cell_dofs()
        ref = oracle.assemble_matrix(oracle.VP2_ELASTICITY, mesh.coordinates(), mesh.cells(),
                                     V.cell_dofs(), rp, col, coeff_dofs=kd, c0=mu, c1=lmbda)
        bref = oracle.assemble_vector(oracle.VP2_BODY_FORCE, mesh.coordinates(), mesh.cells(),
                                      V.cell_dofs(), V.dim(), f3=f3)
        assert np.linalg.norm(vals - ref) / np.linalg.norm(ref) < 1e-12
        assert np.linalg.norm(b.get_local() - bref) / np.linalg.norm(bref) < 1e-12


def test_device_vp2_plan_64(monkeypatch):
    print("case device vp2 plan 64^3", flush=True)
    monkeypatch.setenv("DFA_PLAN_VERIFY", "1")
    mesh, V, mu, lmbda, f3 = workloads.vp2_elasticity_3d(64)
    eng = dfa.fem.Engine(V)
    assert eng.plan_builder() == "device"
