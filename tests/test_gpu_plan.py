"""The device planner of the P1 blocked scatter against the host builder, array by array.

With DFA_PLAN_VERIFY=1 in the environment, creating an assembler builds the P1 block plan,
its edge chains and the chain kernels' slot arrays both on the device (plan_device.hip) and
on the host (structure.cpp) and compares every array element by element; plan_builder()
raises on any difference and names the first differing array. The cases cover the
recursive-bisection ties of an unjittered grid (equal coordinates are ordered by dof), a
stretched box (the split axis changes from level to level), permuted and reverse
Cuthill-McKee dofs (coefficient dof != vertex), dofs that belong to no cell, the block sizes
an autotuner may choose, and the 128^3 benchmark mesh. Without the flag the device plan is
the one assembled with, and every other GPU test runs through it; the assembled values are
checked here against the oracle as well.
"""
import numpy as np
import pytest
import torch

import dolfin_amd as dfa
from dolfin_amd import _lib, autotune, workloads

pytestmark = pytest.mark.gpu


def _builder(V):
    return dfa.engine_for(V).plan_builder()


def _space(n, seed, jitter=0.2, stretch=None, permuted=False, reorder=False):
    rng = np.random.default_rng(seed)
    mesh = dfa.UnitCubeMesh(n, n, n)
    if jitter:
        dfa.jitter_interior(mesh, jitter / n, rng)
    if stretch is not None:
        mesh.set_coordinates(mesh.coordinates() * np.asarray(stretch))
    if permuted:
        return mesh, dfa.FunctionSpace(mesh, "P", 1,
                                       node_permutation=rng.permutation(mesh.num_vertices()))
    return mesh, dfa.FunctionSpace(mesh, "P", 1, reorder=reorder)


CASES = {
    "n2": dict(n=2, seed=1),
    "n3": dict(n=3, seed=2),
    "n8": dict(n=8, seed=3),
    "n17": dict(n=17, seed=4),
    "n8_grid_ties": dict(n=8, seed=5, jitter=0.0),
    "n13_grid_ties": dict(n=13, seed=6, jitter=0.0),
    "n12_stretched": dict(n=12, seed=7, stretch=(4.0, 1.0, 0.25)),
    "n9_permuted": dict(n=9, seed=8, permuted=True),
    "n12_rcm": dict(n=12, seed=9, reorder=True),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_device_plan_equals_host_plan(monkeypatch, name):
    print(f"case device plan {name}", flush=True)
    monkeypatch.setenv("DFA_PLAN_VERIFY", "1")
    mesh, V = _space(**CASES[name])
    assert _builder(V) == "device"


@pytest.mark.parametrize("rows,threads", [(8, 256), (64, 512), (320, 512), (1024, 512)])
def test_device_plan_block_sizes(monkeypatch, rows, threads):
    # blocks whose cell lists outgrow the planner's LDS tables are planned on the host; a
    # device plan, when built, is verified against the host build
    print(f"case device plan rows={rows} threads={threads}", flush=True)
    monkeypatch.setenv("DFA_PLAN_VERIFY", "1")
    autotune.set_plan_defaults(rows, threads, 0)
    try:
        mesh, V = _space(14, seed=10 + rows)
        where = _builder(V)
        print(f"  planned on the {where}", flush=True)
        assert where == "device" or rows > 320
    finally:
        autotune.set_plan_defaults(0, 0, 0)


def test_device_plan_limits_fall_back_to_host(monkeypatch):
    # leaves above 1024 rows are outside the device build: the host builds the plan
    monkeypatch.setenv("DFA_PLAN_VERIFY", "1")
    autotune.set_plan_defaults(2048, 512, 0)
    try:
        mesh, V = _space(14, seed=31)
        assert _builder(V) == "host"
    finally:
        autotune.set_plan_defaults(0, 0, 0)


def test_device_plan_dofs_outside_cells(monkeypatch):
    # a space with dofs no cell holds: the partition covers the held dofs only
    monkeypatch.setenv("DFA_PLAN_VERIFY", "1")

    class Padded:
        element = _lib.ELEMENT_P1

        def __init__(self, mesh, extra):
            self._mesh, self._extra = mesh, extra
            self._W = dfa.FunctionSpace(mesh, "P", 1)

        def mesh(self):
            return self._mesh

        def cell_dofs(self):
            return 2 * self._W.cell_dofs() + 1  # odd dofs only

        def dim(self):
            return 2 * self._mesh.num_vertices() + self._extra

        def coefficient_space(self):
            return self._W

    mesh = dfa.UnitCubeMesh(5, 5, 5)
    V = Padded(mesh, 7)
    eng = dfa.fem.Engine(V)
    assert eng.plan_builder() == "device"


def test_device_plan_assembles_like_the_oracle(oracle):
    mesh, V = _space(8, seed=12, permuted=True)
    assert _builder(V) == "device"
    rng = np.random.default_rng(13)
    kappa, f = rng.uniform(0.5, 2.0, V.dim()), rng.uniform(-1.0, 1.0, V.dim())
    K, F = workloads.to_function(V, kappa), workloads.to_function(V, f)
    A, b = dfa.assemble_system(dfa.poisson_form(V, K), dfa.load_form(V, F))
    torch.cuda.synchronize()
    rp, col, vals = A.data()
    cd = V.cell_dofs()
    ref = oracle.assemble_matrix(oracle.P1_POISSON, mesh.coordinates(), mesh.cells(), cd, rp, col,
                                 coeff_dofs=cd, c0=kappa)
    refb = oracle.assemble_vector(oracle.P1_LOAD, mesh.coordinates(), mesh.cells(), cd, V.dim(),
                                  coeff_dofs=cd, c0=f)
    assert np.linalg.norm(vals - ref) / np.linalg.norm(ref) < 1e-12
    bb = b.values.cpu().numpy()
    assert np.linalg.norm(bb - refb) / np.linalg.norm(refb) < 1e-12


def test_device_plan_128(monkeypatch):
    print("case device plan 128^3", flush=True)
    monkeypatch.setenv("DFA_PLAN_VERIFY", "1")
    mesh, V, kappa, f = workloads.p1_poisson_3d(128)
    eng = dfa.engine_for(V)
    assert eng.plan_builder() == "device"
    assert eng.pass_kernel(_lib.FORM_P1_POISSON, _lib.FORM_P1_LOAD) == \
        "p1_chain_kernel<WANT_B=1,T=512,RR=1>"
