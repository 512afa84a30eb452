"""Reverse Cuthill-McKee node renumbering (DOLFIN's reorder_dofs_serial with the Boost
ordering library). The oracle is pinned by the orderings Boost's own documentation prints
for its cuthill_mckee_ordering example graph (10 vertices, 14 edges: from a given start and
with the automatic pseudo-peripheral start) and by bandwidth properties; the library's C++
ordering must equal the oracle's on graphs whose vertices queue at most 16 new neighbours
at a time (the oracle's stable sort equals libstdc++'s std::sort there). CPU only."""
import ctypes

import numpy as np
import pytest

import dolfin_amd as dfa
from dolfin_amd import _lib

# boost/libs/graph/example/cuthill_mckee_ordering.cpp: edges a-d, a-f, b-c, b-e, b-g, b-j,
# c-d, c-e, d-f, d-i, e-g, f-g, f-h, g-h
BOOST_EDGES = [(0, 3), (0, 5), (1, 2), (1, 4), (1, 6), (1, 9), (2, 3), (2, 4), (3, 5), (3, 8),
               (4, 6), (5, 6), (5, 7), (6, 7)]
# the example's printed reverse Cuthill-McKee orderings (new position -> old vertex)
BOOST_RCM_FROM_6 = [8, 3, 0, 9, 2, 5, 1, 4, 7, 6]
BOOST_RCM_FROM_0 = [9, 1, 4, 6, 7, 2, 8, 5, 3, 0]
BOOST_RCM_AUTO = [0, 8, 5, 7, 3, 6, 4, 2, 1, 9]


def _adj(n, edges):
    adj = [set() for _ in range(n)]
    for a, b in edges:
        adj[a].add(b)
        adj[b].add(a)
    return [sorted(s) for s in adj]


def _csr(adj):
    ptr = np.zeros(len(adj) + 1, np.int32)
    ptr[1:] = np.cumsum([len(a) for a in adj])
    col = np.array([v for a in adj for v in a], np.int32)
    return ptr, col


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _lib_graph_order(adj):
    ptr, col = _csr(adj)
    out = np.empty(len(adj), np.int32)
    col = col if col.size else np.zeros(1, np.int32)
    _lib.check(_lib.load().dfa_graph_reordering(len(adj), _p(ptr), _p(col), _p(out)))
    return out


def _sequence(old_to_new):
    seq = np.empty_like(old_to_new)
    seq[old_to_new] = np.arange(len(old_to_new))
    return seq.tolist()


def test_oracle_matches_boost_example(oracle):
    adj = _adj(10, BOOST_EDGES)
    assert oracle.cuthill_mckee(adj, 6)[::-1] == BOOST_RCM_FROM_6
    assert oracle.cuthill_mckee(adj, 0)[::-1] == BOOST_RCM_FROM_0
    p = oracle.reverse_cuthill_mckee(adj)
    assert _sequence(p) == BOOST_RCM_AUTO
    assert oracle.bandwidth(adj) == 8 and oracle.bandwidth(adj, p) == 4


def test_library_matches_boost_example():
    assert _sequence(_lib_graph_order(_adj(10, BOOST_EDGES))) == BOOST_RCM_AUTO


def test_library_matches_oracle_random_graphs(oracle):
    rng = np.random.default_rng(7)
    for trial in range(30):
        n = int(rng.integers(1, 60))
        m = int(rng.integers(0, 3 * n))
        edges = {(int(a), int(b)) for a, b in rng.integers(0, n, (m, 2)) if a != b}
        adj = _adj(n, edges)
        if max((len(a) for a in adj), default=0) > 16:
            continue
        np.testing.assert_array_equal(_lib_graph_order(adj), oracle.reverse_cuthill_mckee(adj))


@pytest.mark.parametrize("kind", ["square", "cube"])
def test_mesh_node_reordering_matches_oracle(oracle, kind):
    rng = np.random.default_rng(3)
    mesh = dfa.UnitSquareMesh(9, 7) if kind == "square" else dfa.UnitCubeMesh(4, 3, 5)
    nv = mesh.num_vertices()
    # a scrambled vertex numbering, so that the ordering has work to do
    relabel = rng.permutation(nv).astype(np.int32)
    cells = relabel[mesh.cells()]
    got = dfa.node_reordering(cells, nv)
    adj = oracle.node_graph(cells, nv)
    np.testing.assert_array_equal(got, oracle.reverse_cuthill_mckee(adj))
    assert np.array_equal(np.sort(got), np.arange(nv))
    assert oracle.bandwidth(adj, got) < oracle.bandwidth(adj) // 2


def test_reordered_spaces():
    mesh = dfa.UnitCubeMesh(3, 3, 3)
    V = dfa.FunctionSpace(mesh, "P", 1, reorder=True)
    perm = dfa.node_reordering(mesh.cells(), mesh.num_vertices())
    np.testing.assert_array_equal(V.cell_dofs(), perm[mesh.cells()])
    ce, ne = mesh.cell_edges()
    nodes = np.concatenate([mesh.cells(), mesh.num_vertices() + ce], axis=1)
    perm2 = dfa.node_reordering(nodes, mesh.num_vertices() + ne)
    W = dfa.VectorFunctionSpace(mesh, "P", 2, node_permutation=perm2)
    pn = perm2[nodes]
    np.testing.assert_array_equal(W.cell_dofs(), np.concatenate([3 * pn + c for c in range(3)], axis=1))
    # DOLFIN's vector-P2 reordering is not claimed: it has to be passed in
    with pytest.raises(ValueError, match="node_permutation"):
        dfa.VectorFunctionSpace(mesh, "P", 2, reorder=True)
    assert dfa.parameters["reorder_dofs_serial"] is False
    dfa.parameters["reorder_dofs_serial"] = True
    try:
        np.testing.assert_array_equal(dfa.FunctionSpace(mesh, "P", 1).cell_dofs(), V.cell_dofs())
    finally:
        dfa.parameters["reorder_dofs_serial"] = False


def test_reordering_edge_cases():
    # no edges: the identity (BoostGraphOrdering.cpp:50-55); isolated vertices between
    # components; invalid adjacency is rejected
    np.testing.assert_array_equal(_lib_graph_order([[], [], []]), [0, 1, 2])
    adj = _adj(6, [(0, 1), (4, 5)])
    got = _lib_graph_order(adj)
    assert sorted(got.tolist()) == list(range(6))
    ptr = np.array([0, 1, 1], np.int32)
    col = np.array([0], np.int32)  # self loop
    out = np.empty(2, np.int32)
    assert _lib.load().dfa_graph_reordering(2, _p(ptr), _p(col), _p(out)) != 0


def test_oracle_refuses_long_queues(oracle):
    # a star of 18 leaves, started at a leaf (pseudo-peripheral): the centre then queues the
    # 17 other leaves, of equal degree, at once, where Boost's std::sort leaves their order to
    # libstdc++'s introsort, which the oracle does not restate
    adj = _adj(19, [(0, k) for k in range(1, 19)])
    with pytest.raises(ValueError, match="16"):
        oracle.reverse_cuthill_mckee(adj)
    adj17 = _adj(18, [(0, k) for k in range(1, 18)])  # 16 queued at once: still restated
    assert sorted(oracle.reverse_cuthill_mckee(adj17).tolist()) == list(range(18))


def test_kuhn_meshes_stay_in_the_pinned_regime(oracle):
    # the P1 node graph of UnitCubeMesh's Kuhn tetrahedra: at most 14 neighbours per vertex, so
    # no BFS step queues more than 16 and the library's order is the oracle's (Boost's)
    mesh = dfa.UnitCubeMesh(4, 4, 4)
    adj = oracle.node_graph(mesh.cells(), mesh.num_vertices())
    assert max(len(a) for a in adj) == 14
    np.testing.assert_array_equal(dfa.node_reordering(mesh.cells(), mesh.num_vertices()),
                                  oracle.reverse_cuthill_mckee(adj))


def test_vector_sub_dofmaps():
    # python/test/unit/fem/test_dofmap.py:112-120: the vector cell dofs are the component
    # sub-dofmaps concatenated, the components share no dof; and each sub-dofmap gives every
    # node one dof (the same in every cell), distinct nodes distinct dofs
    mesh = dfa.UnitCubeMesh(2, 3, 2)
    ce, ne = mesh.cell_edges()
    nn = mesh.num_vertices() + ne
    nodes = np.concatenate([mesh.cells(), mesh.num_vertices() + ce], axis=1)
    perm = np.random.default_rng(4).permutation(nn)
    for W in (dfa.VectorFunctionSpace(mesh, "P", 2),
              dfa.VectorFunctionSpace(mesh, "P", 2, node_permutation=perm)):
        subs = [W.sub(i).cell_dofs() for i in range(3)]
        np.testing.assert_array_equal(np.concatenate(subs, axis=1), W.cell_dofs())
        for i in range(3):
            for j in range(i + 1, 3):
                assert np.intersect1d(subs[i], subs[j]).size == 0
            dof_of = np.full(nn, -1, np.int64)
            dof_of[nodes.ravel()] = subs[i].ravel()
            np.testing.assert_array_equal(dof_of[nodes], subs[i])      # one dof per node
            assert np.unique(dof_of).size == nn                         # distinct per node
        assert np.array_equal(np.unique(W.cell_dofs()),
                              np.arange(3 * nn))
