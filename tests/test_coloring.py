"""Cell colouring (MeshColoring::color, dolfin/mesh/MeshColoring.cpp:45-189) on the host.

The oracle restates GraphBuilder::local_graph over explicit entities and Boost's
sequential_vertex_coloring. It is pinned by known answers (alternating colours on an
interval mesh, one colour per cell of a fan, all cells of a mesh colour 0 when nothing is
shared) and by the property that fixes a sequential greedy colouring uniquely: neighbours
differ, and a cell of colour k has, for every k' < k, a neighbour of lower index with colour
k'. The library's colouring (dfa_mesh_coloring, shared-vertex counting) must equal the
oracle's bit for bit. The reference's own test (python/test/unit/mesh/test_mesh_coloring.py:
23-33) colours UnitCubeMesh(16) by vertex, edge and facet and checks that it runs.
"""
import numpy as np
import pytest
import scipy.spatial

import dolfin_amd as dfa
from dolfin_amd import MeshColoring


def _greedy_property(graph, colors):
    for c, nbrs in enumerate(graph):
        k = colors[c]
        assert all(colors[n] != k for n in nbrs)
        lower = {colors[n] for n in nbrs if n < c}
        assert set(range(k)) <= lower


def _delaunay(gdim, n, seed):
    pts = np.random.default_rng(seed).random((n, gdim))
    cells = np.sort(scipy.spatial.Delaunay(pts).simplices, axis=1).astype(np.int32)
    return dfa.Mesh(pts, cells)


def test_oracle_known_answers(oracle):
    # interval mesh: consecutive cells share a vertex -> 0, 1, 0, 1, ...
    cells = np.stack([np.arange(9), np.arange(1, 10)], axis=1)
    col, n = oracle.sequential_vertex_coloring(oracle.cell_entity_graph(cells, 0))
    assert n == 2 and col.tolist() == [0, 1] * 4 + [0]
    # triangles around one centre vertex: every pair shares it -> one colour each
    m = 7
    fan = np.array([[0, 1 + i, 1 + (i + 1) % m] for i in range(m)])
    col, n = oracle.sequential_vertex_coloring(oracle.cell_entity_graph(fan, 0))
    assert n == m and col.tolist() == list(range(m))
    # the same fan coloured by edges: cell i shares an edge with i - 1 and i + 1 (odd cycle)
    col, n = oracle.sequential_vertex_coloring(oracle.cell_entity_graph(fan, 1))
    assert n == 3 and col.tolist() == [0, 1, 0, 1, 0, 1, 2]
    # cells share no entity of dimension tdim (only themselves): one colour
    col, n = oracle.sequential_vertex_coloring(oracle.cell_entity_graph(fan, 2))
    assert n == 1 and not col.any()


def test_oracle_greedy_property(oracle):
    mesh = dfa.UnitCubeMesh(3, 3, 3)
    for dim in (0, 1, 2):
        graph = oracle.cell_entity_graph(mesh.cells(), dim)
        col, n = oracle.sequential_vertex_coloring(graph)
        _greedy_property(graph, col)
        assert n == col.max() + 1


@pytest.mark.parametrize("case", ["interval", "square", "cube", "tri300", "tet200"])
@pytest.mark.parametrize("ctype", ["vertex", "edge", "facet"])
def test_library_equals_oracle(oracle, case, ctype):
    mesh = {"interval": lambda: dfa.UnitIntervalMesh(11),
            "square": lambda: dfa.UnitSquareMesh(6, 5),
            "cube": lambda: dfa.UnitCubeMesh(3, 4, 2),
            "tri300": lambda: _delaunay(2, 300, 1),
            "tet200": lambda: _delaunay(3, 200, 2)}[case]()
    dim = MeshColoring.type_to_dim(ctype, mesh)
    if dim > mesh.gdim():
        pytest.skip("no such entity")
    got = MeshColoring.color_cells(mesh, ctype)
    graph = oracle.cell_entity_graph(mesh.cells(), dim)
    want, n = oracle.sequential_vertex_coloring(graph)
    np.testing.assert_array_equal(got, want)
    _greedy_property(graph, got)
    assert len(MeshColoring.entities_of_color(mesh, ctype)) == n


def test_reference_cube_16_types():
    # python/test/unit/mesh/test_mesh_coloring.py:23-33 (UnitCubeMesh(16); vertex, edge, facet)
    mesh = dfa.UnitCubeMesh(16, 16, 16)
    cells = mesh.cells()
    for ctype, shared in (("vertex", 1), ("edge", 2), ("facet", 3)):
        colors = mesh.color(ctype)
        assert colors.shape == (mesh.num_cells(),)
        groups = MeshColoring.entities_of_color(mesh, ctype)
        assert sum(len(g) for g in groups) == mesh.num_cells()
        for g in groups:  # vertex colouring: the cells of one colour have disjoint vertices
            cg = cells[g]
            assert np.all(np.diff(g) > 0)
            if shared == 1:
                assert len(np.unique(cg)) == cg.size
        np.testing.assert_array_equal(MeshColoring.cell_colors(mesh, ctype), colors)
    # vertex colouring needs at least one colour per cell around an interior vertex (24)
    assert len(MeshColoring.entities_of_color(mesh, "vertex")) >= 24
    assert len(MeshColoring.entities_of_color(mesh, "facet")) >= 2


def test_coloring_errors():
    mesh = dfa.UnitSquareMesh(2, 2)
    with pytest.raises(RuntimeError, match="Unknown coloring type"):
        mesh.color("face")
    with pytest.raises(RuntimeError, match="has not been computed"):
        MeshColoring.cell_colors(mesh, "edge")
    with pytest.raises(RuntimeError, match="dim i - j"):
        MeshColoring.color(mesh, (2, 1, 0))
