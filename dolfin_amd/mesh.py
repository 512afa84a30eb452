"""Built-in meshes (host, numpy): the inputs of the assembly path.

Restates DOLFIN's generators so that vertex numbering, cell vertex order and coordinates
are identical:
  UnitIntervalMesh  -> IntervalMesh::build (dolfin/generation/IntervalMesh.cpp)
  UnitSquareMesh    -> RectangleMesh::build_tri, "right"/"left" diagonals
                       (dolfin/generation/RectangleMesh.cpp:46-226)
  UnitCubeMesh      -> BoxMesh::build_tet, six Kuhn tetrahedra per cube
                       (dolfin/generation/BoxMesh.cpp:50-163)
followed by Mesh::order() (cell vertices ascending, dolfin/mesh/MeshEditor.cpp:265-270).
"""
import numpy as np


class Mesh:
    """Simplicial mesh with tdim == gdim: coordinates [nv, gdim], cells [nc, gdim+1] (int32)."""

    def __init__(self, x, cells):
        self._x = np.ascontiguousarray(x, dtype=np.float64)
        if self._x.ndim == 1:
            self._x = self._x.reshape(-1, 1)
        self._cells = np.ascontiguousarray(cells, dtype=np.int32)
        if self._cells.shape[1] != self._x.shape[1] + 1:
            raise ValueError("cells must have gdim+1 vertices (simplices with tdim == gdim)")
        self._edges = None
        self._device = {}
        self._coloring = {}

    def coordinates(self):
        return self._x

    def cells(self):
        return self._cells

    def num_vertices(self):
        return self._x.shape[0]

    def num_cells(self):
        return self._cells.shape[0]

    def gdim(self):
        return self._x.shape[1]

    def hmin(self):
        x = self._x[self._cells]
        n = self._cells.shape[1]
        h = np.inf
        for a in range(n):
            for b in range(a + 1, n):
                h = min(h, float(np.min(np.linalg.norm(x[:, a] - x[:, b], axis=1))))
        return h

    def cell_edges(self):
        """Tet edges numbered by ascending sorted vertex key
        (TopologyComputation::compute_entities_by_key_matching,
        dolfin/mesh/TopologyComputation.cpp:152-330); local edge order of
        TetrahedronCell::create_entities (dolfin/mesh/TetrahedronCell.cpp:119-124)."""
        if self._edges is None:
            if self.gdim() != 3:
                raise ValueError("edges are implemented for tetrahedra")
            loc = np.array([[2, 3], [1, 3], [1, 2], [0, 3], [0, 2], [0, 1]])
            c = self._cells.astype(np.int64)
            a = c[:, loc[:, 0]]
            b = c[:, loc[:, 1]]
            lo = np.minimum(a, b)
            hi = np.maximum(a, b)
            key = lo * self.num_vertices() + hi
            uniq, inv = np.unique(key.ravel(), return_inverse=True)
            self._edges = (inv.reshape(-1, 6).astype(np.int32), int(uniq.shape[0]))
        return self._edges

    def color(self, coloring_type="vertex"):
        """Mesh::color (dolfin/mesh/Mesh.cpp:311-340): cell colours for a type name
        ("vertex", "edge", "facet") or a tuple (tdim, dim, tdim); computed once and kept."""
        return MeshColoring.color(self, MeshColoring.coloring_tuple(self, coloring_type))

    def coordinates_device(self, device="cuda"):
        import torch
        t = self._device.get(device)
        if t is None:
            t = torch.from_numpy(self._x).to(device)
            self._device[device] = t
        return t

    def set_coordinates(self, x):
        self._x = np.ascontiguousarray(x, dtype=np.float64).reshape(self._x.shape)
        self._device = {}


class MeshColoring:
    """Greedy cell colouring, dolfin::MeshColoring (dolfin/mesh/MeshColoring.cpp:45-189).

    Two cells are neighbours when they share an entity of dimension dim (the graph of
    GraphBuilder::local_graph(mesh, tdim, dim), dolfin/graph/GraphBuilder.cpp:124-153); cells
    are visited in index order and each takes the smallest colour no coloured neighbour holds
    (Boost's sequential_vertex_coloring, dolfin/graph/BoostGraphColoring.h:87-113). Computed by
    dfa_mesh_coloring in libdolfin_amd.so (host code; there is no Python fallback).
    """

    @staticmethod
    def type_to_dim(coloring_type, mesh):
        """MeshColoring::type_to_dim (MeshColoring.cpp:170-189)."""
        if coloring_type not in ("vertex", "edge", "facet"):
            raise RuntimeError(f'Unknown coloring type ("{coloring_type}"). Known options are '
                               '"vertex", "edge" and "facet"')
        return {"vertex": 0, "edge": 1}.get(coloring_type, mesh.gdim() - 1)

    @staticmethod
    def coloring_tuple(mesh, coloring_type):
        if isinstance(coloring_type, str):
            d = MeshColoring.type_to_dim(coloring_type, mesh)
            return (mesh.gdim(), d, mesh.gdim())
        t = tuple(int(v) for v in coloring_type)
        if t[0] != t[-1]:
            raise RuntimeError("Mesh coloring does not support dim i - j coloring")
        if len(t) != 3 or t[0] != mesh.gdim() or not 0 <= t[1] <= mesh.gdim():
            raise RuntimeError("cell colourings of type (tdim, dim, tdim) are implemented")
        return t

    @staticmethod
    def compute_colors(mesh, coloring_type):
        """Colours of the cells and their number (MeshColoring::compute_colors, :100-121)."""
        from . import _lib
        import ctypes
        t = MeshColoring.coloring_tuple(mesh, coloring_type)
        cells = np.ascontiguousarray(mesh.cells(), dtype=np.int32)
        colors = np.empty(mesh.num_cells(), dtype=np.int32)
        n = ctypes.c_int32(0)
        _lib.check(_lib.load().dfa_mesh_coloring(mesh.num_cells(), cells.shape[1], _lib.ptr(cells),
                                                 mesh.num_vertices(), t[1], _lib.ptr(colors),
                                                 ctypes.byref(n)))
        return colors.astype(np.int64), n.value

    @staticmethod
    def color(mesh, coloring_type):
        """MeshColoring::color (:57-98): computes, stores on the mesh and returns the colours."""
        t = MeshColoring.coloring_tuple(mesh, coloring_type)
        if t not in mesh._coloring:
            colors, n = MeshColoring.compute_colors(mesh, t)
            order = np.argsort(colors, kind="stable")
            bounds = np.searchsorted(colors[order], np.arange(n + 1))
            mesh._coloring[t] = (colors, [order[bounds[k]:bounds[k + 1]] for k in range(n)])
        return mesh._coloring[t][0]

    @staticmethod
    def color_cells(mesh, coloring_type="vertex"):
        """MeshColoring::color_cells (:45-56)."""
        return MeshColoring.color(mesh, MeshColoring.coloring_tuple(mesh, coloring_type))

    @staticmethod
    def entities_of_color(mesh, coloring_type):
        """Cells of each colour, ascending (the second member of the colouring data kept in
        MeshTopology::coloring, MeshColoring.cpp:88-95)."""
        t = MeshColoring.coloring_tuple(mesh, coloring_type)
        MeshColoring.color(mesh, t)
        return mesh._coloring[t][1]

    @staticmethod
    def cell_colors(mesh, coloring_type):
        """MeshColoring::cell_colors (:126-168): the stored colours; raises if not computed."""
        t = MeshColoring.coloring_tuple(mesh, coloring_type)
        if t not in mesh._coloring:
            raise RuntimeError("Requested coloring has not been computed")
        return mesh._coloring[t][0].copy()


def _lin(n):
    # a + i*(b-a)/n with a = 0, b = 1, evaluated as in the reference
    return (np.arange(n + 1, dtype=np.float64) * (1.0 - 0.0)) / float(n)


def UnitIntervalMesh(n):
    x = _lin(n).reshape(-1, 1)
    i = np.arange(n, dtype=np.int32)
    return Mesh(x, np.stack([i, i + 1], axis=1))


def UnitSquareMesh(nx, ny, diagonal="right"):
    xs, ys = _lin(nx), _lin(ny)
    X = np.empty(((ny + 1) * (nx + 1), 2))
    X[:, 0] = np.tile(xs, ny + 1)
    X[:, 1] = np.repeat(ys, nx + 1)
    iy, ix = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    v0 = (iy * (nx + 1) + ix).ravel().astype(np.int32)
    v1, v2 = v0 + 1, v0 + (nx + 1)
    v3 = v1 + (nx + 1)
    if diagonal == "right":
        t = np.stack([np.stack([v0, v1, v3], 1), np.stack([v0, v2, v3], 1)], axis=1)
    elif diagonal == "left":
        t = np.stack([np.stack([v0, v1, v2], 1), np.stack([v1, v2, v3], 1)], axis=1)
    else:
        raise ValueError("diagonal must be 'right' or 'left'")
    cells = np.sort(t.reshape(-1, 3), axis=1)
    return Mesh(X, cells)


def UnitCubeMesh(nx, ny, nz):
    xs, ys, zs = _lin(nx), _lin(ny), _lin(nz)
    nv = (nx + 1) * (ny + 1) * (nz + 1)
    X = np.empty((nv, 3))
    X[:, 0] = np.tile(xs, (ny + 1) * (nz + 1))
    X[:, 1] = np.tile(np.repeat(ys, nx + 1), nz + 1)
    X[:, 2] = np.repeat(zs, (nx + 1) * (ny + 1))
    iz, iy, ix = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    v0 = (iz * (nx + 1) * (ny + 1) + iy * (nx + 1) + ix).ravel().astype(np.int32)
    sxy = (nx + 1) * (ny + 1)
    v1, v2 = v0 + 1, v0 + (nx + 1)
    v3 = v1 + (nx + 1)
    v4, v5, v6, v7 = v0 + sxy, v1 + sxy, v2 + sxy, v3 + sxy
    t = np.stack([np.stack(c, 1) for c in ((v0, v1, v3, v7), (v0, v1, v7, v5), (v0, v5, v7, v4),
                                           (v0, v3, v2, v7), (v0, v6, v4, v7), (v0, v2, v6, v7))],
                 axis=1)
    cells = np.sort(t.reshape(-1, 4), axis=1)
    return Mesh(X, cells)


def boundary_vertex_mask(mesh, tol=1e-14):
    """Vertices on the boundary of the unit interval/square/cube ("on_boundary")."""
    x = mesh.coordinates()
    return np.any((np.abs(x) < tol) | (np.abs(x - 1.0) < tol), axis=1)


def jitter_interior(mesh, amplitude, rng):
    """Move interior vertices by uniform[-amplitude, amplitude] per coordinate (boundary
    vertices stay), drawing gdim numbers for every vertex in vertex order."""
    x = mesh.coordinates().copy()
    d = rng.uniform(-amplitude, amplitude, size=x.shape)
    inner = ~boundary_vertex_mask(mesh)
    x[inner] += d[inner]
    mesh.set_coordinates(x)
    return mesh
