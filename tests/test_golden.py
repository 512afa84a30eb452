"""Golden fixtures (tests/golden/, written by make_golden.py from the oracle): the oracle
still reproduces them, they satisfy the reference's known answers and the invariants of the
forms, and the package's meshes and dofmaps equal the fixtures' inputs."""
import os
import sys
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import dolfin_amd as dfa
from dolfin_amd import workloads

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
if GOLDEN not in sys.path:
    sys.path.insert(0, GOLDEN)


def load(name):
    with np.load(os.path.join(GOLDEN, name + ".npz")) as z:
        return {k: z[k] for k in z.files}


def read_dolfin_xml(path):
    """Vertices and cells of a DOLFIN XML mesh file (python/test/unit/fem/tetrahedron.xml)."""
    mesh = ET.parse(path).getroot().find("mesh")
    dim = int(mesh.get("dim"))
    verts = sorted(mesh.find("vertices"), key=lambda v: int(v.get("index")))
    x = np.array([[float(v.get(c)) for c in "xyz"[:dim]] for v in verts])
    cells = sorted(mesh.find("cells"), key=lambda c: int(c.get("index")))
    cv = np.array([[int(c.get(f"v{i}")) for i in range(dim + 1)] for c in cells], dtype=np.int32)
    return x, cv


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("name", ["square64_right_p1", "cube4_p1_bc", "cube2_vp2"])
def test_oracle_reproduces_fixture(name):
    import make_golden
    want = load(name)
    got = make_golden.FIXTURES[name]()
    assert set(got) == set(want)
    for k in want:
        if want[k].dtype.kind in "iu":
            np.testing.assert_array_equal(got[k], want[k])
        else:
            assert rel(got[k], want[k]) <= 1e-14 if np.any(want[k]) else not np.any(got[k])


def test_square64_fixture_invariants():
    g = load("square64_right_p1")
    x, cells, rp, col, A = g["x"], g["cells"], g["row_ptr"], g["col"], g["A"]
    assert rp.shape == (4226,) and col.shape == (29057,)  # 4,225 rows, 29,057 stored entries
    rows = np.repeat(np.arange(4225), np.diff(rp))
    # K 1 = 0 and symmetry
    rs = np.zeros(4225)
    np.add.at(rs, rows, A)
    assert np.abs(rs).max() < 1e-13 * np.abs(A).max()
    dense = np.zeros((4225, 4225))
    dense[rows, col] = A
    assert np.abs(dense - dense.T).max() < 1e-15 * np.abs(A).max()
    # sum(b) = integral of the P1 interpolant of f = sum over cells |T| mean(f)
    e1, e2 = x[cells[:, 1]] - x[cells[:, 0]], x[cells[:, 2]] - x[cells[:, 0]]
    area = 0.5 * np.abs(e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0])
    assert abs(g["b"].sum() - np.sum(area * g["f"][cells].mean(axis=1))) < 1e-14
    # the package's mesh and dofmap are the fixture's
    mesh, V, kappa, f = workloads.p1_poisson_2d(64)
    np.testing.assert_array_equal(mesh.coordinates(), x)
    np.testing.assert_array_equal(V.cell_dofs(), cells)
    np.testing.assert_array_equal(kappa, g["kappa"])
    np.testing.assert_array_equal(f, g["f"])


def test_cube4_fixture_known_answer():
    # python/test/unit/fem/test_system_assembler.py:73-91
    g = load("cube4_p1_bc")
    assert round(np.linalg.norm(g["A"]) - 96.847818767384, 10) == 0
    assert round(np.linalg.norm(g["b"]) - 96.564760289080, 10) == 0
    mesh = dfa.UnitCubeMesh(4, 4, 4)
    np.testing.assert_array_equal(mesh.coordinates(), g["x"])
    np.testing.assert_array_equal(mesh.cells(), g["cells"])


def test_cube2_vp2_fixture_invariants():
    g = load("cube2_vp2")
    x, rp, col, A, b = g["x"], g["row_ptr"], g["col"], g["A"], g["b"]
    n = rp.shape[0] - 1
    nn = n // 3
    assert col.shape[0] == 9 * (230 * 8 + 138 * 4 + 24 * 2 + 1)
    # rigid-body modes: K r = 0 for the 3 translations and 3 rotations (P2 nodes: vertices,
    # then edge midpoints; UFC numbering dof = comp * nodes + node)
    mesh = dfa.UnitCubeMesh(2, 2, 2)
    ce, ne = mesh.cell_edges()
    assert nn == x.shape[0] + ne
    pts = np.empty((nn, 3))
    pts[: x.shape[0]] = x
    ecells, eloc = np.unique(ce.ravel(), return_index=True)
    pairs = np.array([[2, 3], [1, 3], [1, 2], [0, 3], [0, 2], [0, 1]])
    cidx, lidx = np.divmod(eloc, 6)
    vv = mesh.cells()[cidx[:, None], pairs[lidx]]
    pts[x.shape[0] + ecells] = 0.5 * (x[vv[:, 0]] + x[vv[:, 1]])
    rows = np.repeat(np.arange(n), np.diff(rp))
    modes = []
    for i in range(3):
        t = np.zeros((3, nn))
        t[i] = 1.0
        modes.append(t.ravel())
    for i, j in ((0, 1), (1, 2), (0, 2)):
        r = np.zeros((3, nn))
        r[i], r[j] = -pts[:, j], pts[:, i]
        modes.append(r.ravel())
    for m in modes:
        y = np.zeros(n)
        np.add.at(y, rows, A * m[col])
        assert np.abs(y).max() < 1e-12 * np.abs(A).max()
    # body force: b_z sums to -|cube|, x and y components vanish
    assert abs(b[2 * nn:].sum() + 1.0) < 1e-14 and not b[: 2 * nn].any()
    V = dfa.VectorFunctionSpace(mesh, "P", 2)
    np.testing.assert_array_equal(V.cell_dofs(), g["cell_dofs"])


def test_tetrahedron_xml_fixture(oracle):
    # python/test/unit/fem/tetrahedron.xml, used by test_assembler.py:351-390
    x, cells = read_dolfin_xml(os.path.join(GOLDEN, "tetrahedron.xml"))
    np.testing.assert_array_equal(x, [[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]])
    np.testing.assert_array_equal(cells, [[0, 1, 2, 3]])
