"""Reverse Cuthill-McKee numbering on the device (dfa_node_reordering / dfa_graph_reordering
with a GPU present): equal to Boost's documented example ordering, to the oracle on random
graphs and on scrambled mesh node graphs, and to the host build (DFA_HOST_RCM=1) on larger
meshes, several components, and the cases the device build hands to the host (a vertex that
queues more than 16 new neighbours at once, more than 16 components)."""
import numpy as np
import pytest

import dolfin_amd as dfa
from test_reorder import BOOST_EDGES, BOOST_RCM_AUTO, _adj, _lib_graph_order, _sequence

pytestmark = pytest.mark.gpu


def test_device_matches_boost_example(monkeypatch):
    monkeypatch.setenv("DFA_RCM_REQUIRE_DEVICE", "1")
    assert _sequence(_lib_graph_order(_adj(10, BOOST_EDGES))) == BOOST_RCM_AUTO


def test_device_matches_oracle_random_graphs(oracle):
    rng = np.random.default_rng(11)
    for trial in range(20):
        n = int(rng.integers(2, 80))
        m = int(rng.integers(1, 3 * n))
        edges = {(int(a), int(b)) for a, b in rng.integers(0, n, (m, 2)) if a != b}
        adj = _adj(n, edges)
        if max((len(a) for a in adj), default=0) > 16:
            continue
        np.testing.assert_array_equal(_lib_graph_order(adj), oracle.reverse_cuthill_mckee(adj))


@pytest.mark.parametrize("kind", ["square", "cube"])
def test_device_mesh_reordering_matches_oracle(oracle, kind):
    rng = np.random.default_rng(3)
    mesh = dfa.UnitSquareMesh(9, 7) if kind == "square" else dfa.UnitCubeMesh(4, 3, 5)
    nv = mesh.num_vertices()
    cells = rng.permutation(nv).astype(np.int32)[mesh.cells()]
    got = dfa.node_reordering(cells, nv)
    np.testing.assert_array_equal(got, oracle.reverse_cuthill_mckee(oracle.node_graph(cells, nv)))


@pytest.mark.parametrize("case", ["cube24", "cube24_scrambled", "p2_nodes", "two_components",
                                  "star", "many_components"])
def test_device_equals_host(case, monkeypatch):
    rng = np.random.default_rng(5)
    if case.startswith("cube24"):
        mesh = dfa.UnitCubeMesh(24, 20, 16)
        n = mesh.num_vertices()
        cells = mesh.cells()
        if case.endswith("scrambled"):
            cells = rng.permutation(n).astype(np.int32)[cells]
    elif case == "p2_nodes":
        # vector-P2 node graph: vertices queue more than 16 neighbours (the host build runs)
        mesh = dfa.UnitCubeMesh(4, 4, 4)
        ce, ne = mesh.cell_edges()
        cells = np.ascontiguousarray(np.concatenate([mesh.cells(), mesh.num_vertices() + ce], axis=1),
                                     dtype=np.int32)
        n = mesh.num_vertices() + ne
    elif case == "two_components":
        m1, m2 = dfa.UnitCubeMesh(5, 4, 3), dfa.UnitSquareMesh(6, 5)
        c2 = np.pad(m2.cells(), ((0, 0), (0, 1)), mode="edge") + m1.num_vertices()
        cells = np.ascontiguousarray(np.concatenate([m1.cells(), c2]), dtype=np.int32)
        n = m1.num_vertices() + m2.num_vertices()
    elif case == "star":
        # a centre with 18 leaves (line cells), the centre queues 17 leaves at once
        cells = np.array([[0, k] for k in range(1, 19)], np.int32)
        n = 19
    else:
        # 40 disjoint segments: more components than the device build takes
        cells = np.arange(80, dtype=np.int32).reshape(40, 2)
        n = 80
    if case in ("cube24", "cube24_scrambled", "two_components"):
        # these run on the device (DFA_RCM_REQUIRE_DEVICE makes a fallback an error); the
        # others go to the host build by design
        monkeypatch.setenv("DFA_RCM_REQUIRE_DEVICE", "1")
    dev = dfa.node_reordering(cells, n)
    monkeypatch.delenv("DFA_RCM_REQUIRE_DEVICE", raising=False)
    monkeypatch.setenv("DFA_HOST_RCM", "1")
    host = dfa.node_reordering(cells, n)
    np.testing.assert_array_equal(dev, host)
    assert np.array_equal(np.sort(dev), np.arange(n))
