"""GPU parity against the committed golden fixtures (tests/golden/): the HIP path, through
the C ABI, on the fixtures' inputs. Bar: CSR structure bit-exact; values within relative
Frobenius / L2 error 1e-12 in FP64 (summation order differs from the serial oracle)."""
import os

import numpy as np
import pytest
import torch

import dolfin_amd as dfa
from dolfin_amd import workloads
from test_golden import GOLDEN, load, read_dolfin_xml

pytestmark = pytest.mark.gpu

TOL = 1e-12
P1_SCATTERS = [dfa.SCATTER_ATOMIC, dfa.SCATTER_BLOCKED, dfa.SCATTER_COLORED]


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("scatter", P1_SCATTERS)
def test_square64_right_p1(scatter):
    print(f"case golden square64 scatter={scatter}", flush=True)
    g = load("square64_right_p1")
    mesh = dfa.Mesh(g["x"], g["cells"])
    V = dfa.FunctionSpace(mesh, "P", 1)
    A = dfa.assemble(dfa.poisson_form(V, workloads.to_function(V, g["kappa"])), scatter=scatter)
    b = dfa.assemble(dfa.load_form(V, workloads.to_function(V, g["f"])), scatter=scatter)
    rp, col, vals = A.data()
    np.testing.assert_array_equal(rp, g["row_ptr"])
    np.testing.assert_array_equal(col, g["col"])
    assert rel(vals, g["A"]) < TOL
    assert rel(b.get_local(), g["b"]) < TOL


@pytest.mark.parametrize("scatter", P1_SCATTERS)
def test_cube4_p1_bc(scatter):
    print(f"case golden cube4 bc scatter={scatter}", flush=True)
    g = load("cube4_p1_bc")
    mesh = dfa.Mesh(g["x"], g["cells"])
    V = dfa.FunctionSpace(mesh, "P", 1)
    one = dfa.Function(V, torch.ones(V.dim(), dtype=torch.float64))
    ten = dfa.Function(V, torch.full((V.dim(),), 10.0, dtype=torch.float64))
    bc = dfa.DirichletBC(V, 1.0, "on_boundary")
    A, b = dfa.assemble_system(dfa.poisson_form(V, one), dfa.load_form(V, ten), bc,
                               scatter=scatter)
    rp, col, vals = A.data()
    np.testing.assert_array_equal(rp, g["row_ptr"])
    np.testing.assert_array_equal(col, g["col"])
    assert rel(vals, g["A"]) < TOL
    assert rel(b.get_local(), g["b"]) < TOL


@pytest.mark.parametrize("scatter", [dfa.SCATTER_ATOMIC, dfa.SCATTER_BLOCKED])
def test_cube2_vp2(scatter):
    print(f"case golden cube2 vp2 scatter={scatter}", flush=True)
    g = load("cube2_vp2")
    mesh = dfa.Mesh(g["x"], g["cells"])
    V = dfa.VectorFunctionSpace(mesh, "P", 2)
    np.testing.assert_array_equal(V.cell_dofs(), g["cell_dofs"])
    Q = V.coefficient_space()
    A = dfa.assemble(dfa.elasticity_form(V, workloads.to_function(Q, g["mu"]),
                                         workloads.to_function(Q, g["lmbda"])), scatter=scatter)
    b = dfa.assemble(dfa.body_force_form(V, g["f3"]), scatter=scatter)
    rp, col, vals = A.data()
    np.testing.assert_array_equal(rp, g["row_ptr"])
    np.testing.assert_array_equal(col, g["col"])
    assert rel(vals, g["A"]) < TOL
    assert rel(b.get_local(), g["b"]) < TOL


@pytest.mark.parametrize("scatter", P1_SCATTERS)
def test_reference_tetrahedron_xml(scatter):
    # python/test/unit/fem/test_assembler.py:351-390 on the reference's tetrahedron.xml
    print(f"case golden tetrahedron.xml scatter={scatter}", flush=True)
    x, cells = read_dolfin_xml(os.path.join(GOLDEN, "tetrahedron.xml"))
    V = dfa.FunctionSpace(dfa.Mesh(x, cells), "P", 1)
    one = dfa.Function(V, torch.ones(V.dim(), dtype=torch.float64))
    A = dfa.assemble(dfa.poisson_form(V, one), scatter=scatter)
    M = dfa.assemble(dfa.mass_form(V), scatter=scatter)
    A0 = np.array([[1.0 / 2.0, -1.0 / 6.0, -1.0 / 6.0, -1.0 / 6.0],
                   [-1.0 / 6.0, 1.0 / 6.0, 0.0, 0.0],
                   [-1.0 / 6.0, 0.0, 1.0 / 6.0, 0.0],
                   [-1.0 / 6.0, 0.0, 0.0, 1.0 / 6.0]])
    M0 = np.full((4, 4), 1.0 / 120.0) + np.eye(4) / 120.0
    for T, T0 in ((A, A0), (M, M0)):
        rp, col, vals = T.data()
        dense = np.zeros((4, 4))
        dense[np.repeat(np.arange(4), np.diff(rp)), col] = vals
        assert np.linalg.norm(dense - T0) < 1e-14
