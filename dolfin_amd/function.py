"""Function spaces, dofmaps and coefficient functions (host construction, device storage).

Dof numbering (the input that fixes the CSR structure):
  P1            : UFC numbering, dof = vertex (DofMapBuilder with reorder_dofs_serial=False,
                  dolfin/fem/DofMapBuilder.cpp:253-272), or dof = perm[vertex] when a node
                  permutation (e.g. DOLFIN's Boost RCM order) is supplied.
  vector P2     : nodes are vertices then edges (node = nv + edge); UFC numbering is
                  component-major, dof = comp*(nv+ne) + node; with a node permutation the
                  numbering is block-interleaved, dof = 3*perm[node] + comp
                  (DofMapBuilder::build_dofmap, dolfin/fem/DofMapBuilder.cpp:1669-1693).
  Cell-local order is component-major [comp][4 vertices, 6 edges]
  (python/test/unit/fem/test_dofmap.py:112-120).
"""
import numpy as np
import torch

from . import _lib


# Global parameters read at FunctionSpace construction (the dof-ordering subset of
# dolfin::parameters, dolfin/parameter/GlobalParameters.h:76-88). reorder_dofs_serial
# defaults to False here (DOLFIN: True) so that P1 dofs are vertex indices.
parameters = {"reorder_dofs_serial": False}


def node_reordering(cell_nodes, num_nodes):
    """old_to_new of DOLFIN's reverse Cuthill-McKee node renumbering (DofMapBuilder::
    compute_node_reordering, DofMapBuilder.cpp:1473-1569, Boost ordering library). Equal to
    Boost's order when no vertex queues more than 16 new neighbours at once (P1 on the Kuhn
    meshes of UnitCubeMesh: at most 14 neighbours); beyond that, equal-degree neighbours are
    ordered by libstdc++'s std::sort, as a GCC build of Boost orders them, with no test to pin
    it."""
    cn = np.ascontiguousarray(cell_nodes, dtype=np.int32)
    out = np.empty(int(num_nodes), dtype=np.int32)
    _lib.check(_lib.load().dfa_node_reordering(cn.shape[0], cn.shape[1], _lib.ptr(cn),
                                               int(num_nodes), _lib.ptr(out)))
    return out


class FunctionSpace:
    def __init__(self, mesh, family="P", degree=1, vector=False, node_permutation=None,
                 reorder=None):
        """node_permutation: new index of each node (vertex for P1; vertex, then nv + edge
        for vector P2), or reorder=True for DOLFIN's reverse Cuthill-McKee numbering
        (default: parameters["reorder_dofs_serial"]). A permuted vector space numbers its
        dofs 3 * node + component, as DOLFIN does after reordering (DofMapBuilder.cpp:1689)."""
        if family not in ("P", "Lagrange", "CG"):
            raise ValueError("only continuous Lagrange spaces are supported")
        if reorder is None:
            reorder = parameters["reorder_dofs_serial"]
        if reorder and node_permutation is not None:
            raise ValueError("give either node_permutation or reorder, not both")
        self._mesh = mesh
        self.degree = degree
        self.vector = vector
        nc = mesh.num_cells()
        if degree == 1 and not vector:
            self.element = _lib.ELEMENT_P1
            nv = mesh.num_vertices()
            if reorder:
                node_permutation = node_reordering(mesh.cells(), nv)
            if node_permutation is None:
                cd = mesh.cells().copy()
            else:
                perm = np.asarray(node_permutation, dtype=np.int32)
                if perm.shape != (nv,) or not np.array_equal(np.sort(perm), np.arange(nv)):
                    raise ValueError("node_permutation must permute the vertices")
                cd = perm[mesh.cells()]
            self._dim = nv
            self._cell_dofs = np.ascontiguousarray(cd, dtype=np.int32)
            self._node_coords = mesh.coordinates()
        elif degree == 2 and vector:
            if mesh.gdim() != 3:
                raise ValueError("vector P2 is implemented on tetrahedra")
            self.element = _lib.ELEMENT_VECTOR_P2
            ce, ne = mesh.cell_edges()
            nv = mesh.num_vertices()
            nn = nv + ne
            nodes = np.concatenate([mesh.cells(), nv + ce], axis=1).astype(np.int64)  # [nc, 10]
            if reorder:
                # DOLFIN orders these nodes with Boost's reverse Cuthill-McKee, whose order of
                # equal-degree neighbours for the long BFS queues of the P2 node graph (a node
                # has up to ~60 neighbours) is not reproduced bit for bit here; DOLFIN's own
                # permutation is passed in instead (its DofMap gives it)
                raise ValueError("reorder_dofs_serial for vector P2: pass DOLFIN's node "
                                 "permutation as node_permutation (this build's reverse "
                                 "Cuthill-McKee is node_reordering(), a locality ordering "
                                 "with no claim to equal Boost's for queues over 16)")
            if node_permutation is None:
                cd = np.concatenate([c * nn + nodes for c in range(3)], axis=1)
            else:
                perm = np.asarray(node_permutation, dtype=np.int64)
                if perm.shape != (nn,):
                    raise ValueError("node_permutation must permute the P2 nodes")
                pn = perm[nodes]
                cd = np.concatenate([3 * pn + c for c in range(3)], axis=1)
            self._dim = 3 * nn
            self._cell_dofs = np.ascontiguousarray(cd, dtype=np.int32)
        else:
            raise ValueError("supported spaces: FunctionSpace(mesh,'P',1), VectorFunctionSpace(mesh,'P',2)")
        self._engines = {}
        self._p1 = None

    def mesh(self):
        return self._mesh

    def dim(self):
        return self._dim

    def cell_dofs(self):
        return self._cell_dofs

    def sub(self, i):
        """Component i of a vector space (FunctionSpace::sub): its dofmap is the columns
        i*10 .. i*10 + 9 of the component-major cell dofs (python/test/unit/fem/
        test_dofmap.py:112-120)."""
        if self.element != _lib.ELEMENT_VECTOR_P2:
            raise ValueError("sub() needs a vector space")
        if not 0 <= i < 3:
            raise IndexError("component index out of range")
        return _SubSpace(self, i)

    def coefficient_space(self):
        """P1 space whose dofmap addresses the P1 coefficients of forms on this space."""
        if self.element == _lib.ELEMENT_P1:
            return self
        if self._p1 is None:
            self._p1 = FunctionSpace(self._mesh, "P", 1)
        return self._p1

    def __repr__(self):
        kind = "VectorP2" if self.vector else "P1"
        return f"FunctionSpace({kind}, dim={self._dim})"


class _SubSpace:
    """A component of a vector P2 space: dofs of one component, numbered in the parent."""

    def __init__(self, parent, i):
        self._parent, self._i = parent, i

    def mesh(self):
        return self._parent.mesh()

    def cell_dofs(self):
        return self._parent.cell_dofs()[:, 10 * self._i:10 * self._i + 10]


def VectorFunctionSpace(mesh, family="P", degree=2, node_permutation=None, reorder=None):
    return FunctionSpace(mesh, family, degree, vector=True, node_permutation=node_permutation,
                         reorder=reorder)


class Function:
    """Finite element function with its coefficient vector resident on the device."""

    def __init__(self, V, values=None, device="cuda"):
        self._V = V
        if values is None:
            self._vec = torch.zeros(V.dim(), dtype=torch.float64, device=device)
        else:
            v = torch.as_tensor(values, dtype=torch.float64)
            if v.shape != (V.dim(),):
                raise ValueError(f"expected {V.dim()} values, got {tuple(v.shape)}")
            self._vec = v.to(device).contiguous()

    def function_space(self):
        return self._V

    def vector(self):
        return self._vec

    def assign(self, values):
        self._vec.copy_(torch.as_tensor(values, dtype=torch.float64))
