"""The graph-coloured scatter's colouring on the device (Jones-Plassmann over the cells that
share a dof): a valid colouring for the scatter, checked through the assembled values against
the oracle and against the same scatter with DOLFIN's sequential greedy colouring on the host
(DFA_HOST_COLOR=1), on structured and Delaunay meshes."""
import numpy as np
import pytest
import scipy.spatial

import dolfin_amd as dfa
from dolfin_amd import workloads

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _mesh(kind, seed):
    rng = np.random.default_rng(seed)
    if kind == "cube":
        mesh = dfa.UnitCubeMesh(9, 7, 8)
        workloads.jitter_interior(mesh, 0.2 / 9, rng)
        return mesh, rng
    gdim = 2 if kind == "delaunay2" else 3
    corners = np.array(np.meshgrid(*[[0.0, 1.0]] * gdim, indexing="ij")).reshape(gdim, -1).T
    pts = np.vstack([rng.uniform(0, 1, (600, gdim)), corners])
    cells = np.sort(scipy.spatial.Delaunay(pts).simplices, axis=1).astype(np.int32)
    return dfa.Mesh(pts, cells), rng


@pytest.mark.parametrize("kind", ["cube", "delaunay2", "delaunay3"])
def test_device_coloring_scatter(oracle, kind, monkeypatch):
    print(f"case coloring {kind}", flush=True)
    mesh, rng = _mesh(kind, seed=11)
    results = {}
    for where in ("device", "host"):
        if where == "host":
            monkeypatch.setenv("DFA_HOST_COLOR", "1")
        V = dfa.FunctionSpace(mesh, "P", 1)   # a fresh space: its own handle and colour plan
        kappa = np.random.default_rng(1).uniform(0.5, 2, V.dim())
        f = np.random.default_rng(2).uniform(-1, 1, V.dim())
        A, b = dfa.assemble_system(dfa.poisson_form(V, workloads.to_function(V, kappa)),
                                   dfa.load_form(V, workloads.to_function(V, f)),
                                   scatter=dfa.SCATTER_COLORED)
        ptr = dfa.engine_for(V).color_ptr()
        results[where] = (A.data(), b.get_local(), ptr, V, kappa, f)
    (rp, col, vals), bv, ptr, V, kappa, f = results["device"]
    ref = oracle.assemble_matrix(oracle.P1_POISSON, mesh.coordinates(), mesh.cells(), V.cell_dofs(),
                                 rp, col, c0=kappa)
    bref = oracle.assemble_vector(oracle.P1_LOAD, mesh.coordinates(), mesh.cells(), V.cell_dofs(),
                                  V.dim(), c0=f)
    assert _rel(vals, ref) < TOL and _rel(bv, bref) < TOL
    (_, _, hvals), hb, hptr, _, _, _ = results["host"]
    assert _rel(hvals, ref) < TOL and _rel(hb, bref) < TOL
    # every cell in exactly one colour; the parallel colouring needs at most twice the colours
    # of the sequential greedy one
    assert ptr[0] == 0 and ptr[-1] == mesh.num_cells() and np.all(np.diff(ptr) > 0)
    assert len(ptr) - 1 <= 2 * (len(hptr) - 1)
