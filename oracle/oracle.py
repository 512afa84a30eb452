"""ctypes front-end of the C++ oracle (oracle.cpp) -- test infrastructure only.

Every function here is a thin numpy wrapper; the arithmetic lives in oracle.cpp, which
cites the DOLFIN 2019.1 source lines it restates.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "_build", "liboracle.so")

(P1_POISSON, P1_LOAD, P1_MASS, VP2_ELASTICITY, VP2_BODY_FORCE, VP1_ELASTICITY, VP1_BODY_FORCE,
 P1_FACET_MASS, P1_FACET_LOAD, VP2_FACET_TRACTION) = range(10)

_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        _lib = ctypes.CDLL(_SO)
        _declare(_lib)
    return _lib


def use_native_build():
    """Rebuild the oracle with -O3 -march=native for the host this runs on and route every
    later call of this module through that build (the CPU-baseline timing uses it). Returns
    the compiler flags used."""
    global _lib
    subprocess.check_call(["make", "-s", "-C", _HERE, "native"])
    _lib = ctypes.CDLL(os.path.join(_HERE, "_build", "liboracle_native.so"))
    _declare(_lib)
    return "-O3 -march=native"


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo), or "" where that file is absent."""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return ""


_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_i64 = ctypes.c_int64
_i32 = ctypes.c_int32
_vp = ctypes.c_void_p


def _declare(L):
    L.orc_interval_mesh.argtypes = [_i32, _f64p, _i32p]
    L.orc_rectangle_mesh_right.argtypes = [_i32, _i32, _f64p, _i32p]
    L.orc_box_mesh.argtypes = [_i32, _i32, _i32, _f64p, _i32p]
    L.orc_tet_edges.argtypes = [_i64, _i32p, _i32p]
    L.orc_tet_edges.restype = _i64
    L.orc_vp2_dofmap.argtypes = [_i64, _i32p, _i32p, _i64, _i64, _i32p]
    L.orc_sparsity_build.argtypes = [_i64, _i32, _i32p, _i64, _i32]
    L.orc_sparsity_build.restype = _vp
    L.orc_sparsity_nnz.argtypes = [_vp]
    L.orc_sparsity_nnz.restype = _i64
    L.orc_sparsity_copy.argtypes = [_vp, _i32p, _i32p]
    L.orc_sparsity_free.argtypes = [_vp]
    for name in ("orc_p1_stiffness",):
        getattr(L, name).argtypes = [_i32, _f64p, _f64p, _f64p]
    L.orc_p1_mass.argtypes = [_i32, _f64p, _f64p]
    L.orc_p1_load.argtypes = [_i32, _f64p, _f64p, _f64p]
    L.orc_vp2_elasticity.argtypes = [_f64p, _f64p, _f64p, _f64p]
    L.orc_vp1_elasticity.argtypes = [_f64p, _f64p, _f64p, _f64p]
    L.orc_vp2_body_force.argtypes = [_f64p, _f64p, _f64p]
    L.orc_vp1_body_force.argtypes = [_f64p, _f64p, _f64p]
    L.orc_assemble_matrix.argtypes = [_i32, _i32, _i64, _i32p, _f64p, _i32, _i32p, _i32p,
                                      _vp, _vp, _i32p, _i32p, _f64p]
    L.orc_assemble_matrix.restype = _i32
    L.orc_assemble_vector.argtypes = [_i32, _i32, _i64, _i32p, _f64p, _i32, _i32p, _i32p,
                                      _vp, _vp, _f64p]
    L.orc_assemble_vector.restype = _i32
    L.orc_assemble_system.argtypes = [_i32, _i32, _i32, _i64, _i32p, _f64p, _i32, _i32p, _i32p,
                                      _vp, _vp, _vp, _vp, _u8p, _f64p, _i32p, _i32p, _vp, _vp]
    L.orc_assemble_system.restype = _i32
    L.orc_assemble_system_threads.argtypes = L.orc_assemble_system.argtypes + [_i32]
    L.orc_assemble_system_threads.restype = _i32
    L.orc_exterior_facets.argtypes = [_i32, _i64, _i32p, _vp, _vp]
    L.orc_exterior_facets.restype = _i64
    L.orc_p1_facet_mass.argtypes = [_i32, _f64p, _i32, ctypes.c_double, _f64p]
    L.orc_p1_facet_load.argtypes = [_i32, _f64p, _i32, _f64p, _f64p]
    L.orc_vp2_facet_traction.argtypes = [_f64p, _i32, _f64p, _f64p]
    L.orc_assemble_facets.argtypes = [_i32, _i32, _i32, _i32p, _f64p, _i32p, _i32p, _i64, _i32p,
                                      _i32p, _vp, ctypes.c_double, _vp, _i32p, _i32p, _vp, _vp]
    L.orc_assemble_facets.restype = _i32
    L.orc_assemble_system_facets.argtypes = [_i32, _i32, _i32, _i32, _i32, _i64, _i32p, _f64p,
                                             _i32p, _i32p, _vp, _vp, ctypes.c_double, _vp, _i64,
                                             _i32p, _i32p, _vp, _u8p, _f64p, _i32p, _i32p, _vp,
                                             _vp]
    L.orc_assemble_system_facets.restype = _i32
    L.orc_adjoint_p1.argtypes = [_i32, _i64, _i32p, _f64p, _i64, _i32p, _i64, _f64p, _f64p,
                                 _f64p, _f64p, _f64p]
    L.orc_adjoint_p1.restype = ctypes.c_double
    L.orc_adjoint_vp2.argtypes = [_i64, _i32p, _f64p, _i64, _i32p, _i32p, _i64, _f64p, _f64p,
                                  _f64p, _f64p, _f64p, _f64p, _f64p]
    L.orc_adjoint_vp2.restype = ctypes.c_double


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


# ---- meshes -------------------------------------------------------------------------
def interval_mesh(n):
    x = np.empty(n + 1)
    cells = np.empty((n, 2), np.int32)
    lib().orc_interval_mesh(n, x, cells)
    return x.reshape(-1, 1), cells


def rectangle_mesh_right(nx, ny):
    x = np.empty(((nx + 1) * (ny + 1), 2))
    cells = np.empty((2 * nx * ny, 3), np.int32)
    lib().orc_rectangle_mesh_right(nx, ny, x, cells)
    return x, cells


def box_mesh(nx, ny, nz):
    x = np.empty(((nx + 1) * (ny + 1) * (nz + 1), 3))
    cells = np.empty((6 * nx * ny * nz, 4), np.int32)
    lib().orc_box_mesh(nx, ny, nz, x, cells)
    return x, cells


def tet_edges(cells):
    cells = _c(cells, np.int32)
    ce = np.empty((cells.shape[0], 6), np.int32)
    ne = lib().orc_tet_edges(cells.shape[0], cells, ce)
    return ce, int(ne)


def vp2_dofmap(cells, cell_edges, nv, ne):
    cells = _c(cells, np.int32)
    cd = np.empty((cells.shape[0], 30), np.int32)
    lib().orc_vp2_dofmap(cells.shape[0], cells, _c(cell_edges, np.int32), nv, ne, cd)
    return cd


def sparsity(cell_dofs, num_rows, keep_diagonal=False):
    cd = _c(cell_dofs, np.int32)
    h = lib().orc_sparsity_build(cd.shape[0], cd.shape[1], cd, num_rows, int(keep_diagonal))
    nnz = lib().orc_sparsity_nnz(h)
    rp = np.empty(num_rows + 1, np.int32)
    col = np.empty(nnz, np.int32)
    lib().orc_sparsity_copy(h, rp, col)
    lib().orc_sparsity_free(h)
    return rp, col


# ---- element tensors ---------------------------------------------------------------
def p1_stiffness(xc, kappa):
    xc = _c(xc, np.float64)
    g = xc.shape[1]
    A = np.empty((g + 1) ** 2)
    lib().orc_p1_stiffness(g, xc.ravel(), _c(kappa, np.float64), A)
    return A.reshape(g + 1, g + 1)


def p1_mass(xc):
    xc = _c(xc, np.float64)
    g = xc.shape[1]
    A = np.empty((g + 1) ** 2)
    lib().orc_p1_mass(g, xc.ravel(), A)
    return A.reshape(g + 1, g + 1)


def p1_load(xc, f):
    xc = _c(xc, np.float64)
    g = xc.shape[1]
    b = np.empty(g + 1)
    lib().orc_p1_load(g, xc.ravel(), _c(f, np.float64), b)
    return b


def vp2_elasticity(xc, mu, lmbda):
    A = np.empty(900)
    lib().orc_vp2_elasticity(_c(xc, np.float64).ravel(), _c(mu, np.float64), _c(lmbda, np.float64), A)
    return A.reshape(30, 30)


def vp1_elasticity(xc, mu, lmbda):
    A = np.empty(144)
    lib().orc_vp1_elasticity(_c(xc, np.float64).ravel(), _c(mu, np.float64), _c(lmbda, np.float64), A)
    return A.reshape(12, 12)


def vp2_body_force(xc, f3):
    b = np.empty(30)
    lib().orc_vp2_body_force(_c(xc, np.float64).ravel(), _c(f3, np.float64), b)
    return b


def vp1_body_force(xc, f3):
    b = np.empty(12)
    lib().orc_vp1_body_force(_c(xc, np.float64).ravel(), _c(f3, np.float64), b)
    return b


# ---- assembly ---------------------------------------------------------------------
def assemble_matrix(kind, x, cells, cell_dofs, row_ptr, col, coeff_dofs=None, c0=None, c1=None):
    x = _c(x, np.float64)
    cells = _c(cells, np.int32)
    cd = _c(cell_dofs, np.int32)
    kd = _c(cells if coeff_dofs is None else coeff_dofs, np.int32)
    c0 = None if c0 is None else _c(c0, np.float64)
    c1 = None if c1 is None else _c(c1, np.float64)
    vals = np.zeros(col.shape[0])
    rc = lib().orc_assemble_matrix(kind, x.shape[1], cells.shape[0], cells, x.ravel(), cd.shape[1],
                                   cd, kd, _ptr(c0), _ptr(c1), _c(row_ptr, np.int32),
                                   _c(col, np.int32), vals)
    if rc != 0:
        raise RuntimeError(f"oracle assemble_matrix failed ({rc})")
    return vals


def assemble_vector(kind, x, cells, cell_dofs, ndofs, coeff_dofs=None, c0=None, f3=None):
    x = _c(x, np.float64)
    cells = _c(cells, np.int32)
    cd = _c(cell_dofs, np.int32)
    kd = _c(cells if coeff_dofs is None else coeff_dofs, np.int32)
    c0 = None if c0 is None else _c(c0, np.float64)
    f3 = None if f3 is None else _c(f3, np.float64)
    b = np.zeros(ndofs)
    rc = lib().orc_assemble_vector(kind, x.shape[1], cells.shape[0], cells, x.ravel(), cd.shape[1],
                                   cd, kd, _ptr(c0), _ptr(f3), b)
    if rc != 0:
        raise RuntimeError(f"oracle assemble_vector failed ({rc})")
    return b


def assemble_system(a_kind, L_kind, x, cells, cell_dofs, row_ptr, col, bc_mask, bc_value,
                    coeff_dofs=None, a_c0=None, a_c1=None, L_c0=None, f3=None,
                    want_A=True, want_b=True, threads=0, x0=None):
    """threads=0: the serial loop (ascending cells); threads>0: the same loop on that many host
    threads with atomic scatter (for CPU timing; equal up to summation order).
    x0: SystemAssembler::assemble(A, b, x0), dolfin/fem/SystemAssembler.cpp:252-275 -- before
    the cell loop every boundary value g of a Dirichlet dof becomes x0[dof] - g."""
    if x0 is not None:
        m = np.asarray(bc_mask) != 0
        bc_value = np.where(m, np.asarray(x0, dtype=np.float64) - np.asarray(bc_value, dtype=np.float64), 0.0)
    x = _c(x, np.float64)
    cells = _c(cells, np.int32)
    cd = _c(cell_dofs, np.int32)
    kd = _c(cells if coeff_dofs is None else coeff_dofs, np.int32)
    arrs = [None if a is None else _c(a, np.float64) for a in (a_c0, a_c1, L_c0, f3)]
    ndofs = row_ptr.shape[0] - 1
    vals = np.zeros(col.shape[0]) if want_A else None
    b = np.zeros(ndofs) if want_b else None
    args = (a_kind, L_kind, x.shape[1], cells.shape[0], cells, x.ravel(), cd.shape[1], cd, kd,
            *[_ptr(a) for a in arrs], _c(bc_mask, np.uint8), _c(bc_value, np.float64),
            _c(row_ptr, np.int32), _c(col, np.int32), _ptr(vals), _ptr(b))
    if threads > 0:
        rc = lib().orc_assemble_system_threads(*args, int(threads))
    else:
        rc = lib().orc_assemble_system(*args)
    if rc != 0:
        raise RuntimeError(f"oracle assemble_system failed ({rc})")
    return vals, b


# ---- exterior facets -------------------------------------------------------------------
def exterior_facets(cells):
    """(facet_cell, facet_local) of the facets with one incident cell."""
    cells = _c(cells, np.int32)
    gdim = cells.shape[1] - 1
    n = lib().orc_exterior_facets(gdim, cells.shape[0], cells, None, None)
    fc = np.empty(n, np.int32)
    fl = np.empty(n, np.int32)
    lib().orc_exterior_facets(gdim, cells.shape[0], cells, _ptr(fc), _ptr(fl))
    return fc, fl


def p1_facet_mass(xc, local_facet, alpha=1.0):
    xc = _c(xc, np.float64)
    gdim = xc.shape[1]
    A = np.empty((gdim + 1) ** 2)
    lib().orc_p1_facet_mass(gdim, xc.ravel(), int(local_facet), float(alpha), A)
    return A.reshape(gdim + 1, gdim + 1)


def p1_facet_load(xc, local_facet, g):
    xc = _c(xc, np.float64)
    gdim = xc.shape[1]
    b = np.empty(gdim + 1)
    lib().orc_p1_facet_load(gdim, xc.ravel(), int(local_facet), _c(g, np.float64), b)
    return b


def vp2_facet_traction(xc, local_facet, t3):
    """inner(t, v)*ds on one facet of a tetrahedron, vector-P2 component-major (30 entries)."""
    b = np.empty(30)
    lib().orc_vp2_facet_traction(_c(xc, np.float64).ravel(), int(local_facet), _c(t3, np.float64), b)
    return b


def assemble_facets(a_kind, L_kind, x, cells, cell_dofs, facets, row_ptr=None, col=None,
                    ndofs=None, coeff_dofs=None, alpha=1.0, g=None, facet_mask=None):
    """Exterior-facet terms (Assembler semantics): returns (values or None, b or None)."""
    x = _c(x, np.float64)
    cells = _c(cells, np.int32)
    cd = _c(cell_dofs, np.int32)
    kd = _c(cells if coeff_dofs is None else coeff_dofs, np.int32)
    fc, fl = (_c(a, np.int32) for a in facets)
    mask = None if facet_mask is None else _c(facet_mask, np.uint8)
    g = None if g is None else _c(g, np.float64)
    vals = np.zeros(col.shape[0]) if a_kind >= 0 else None
    b = np.zeros(ndofs if ndofs is not None else row_ptr.shape[0] - 1) if L_kind >= 0 else None
    rp = _c(row_ptr, np.int32) if row_ptr is not None else np.zeros(1, np.int32)
    cl = _c(col, np.int32) if col is not None else np.zeros(1, np.int32)
    rc = lib().orc_assemble_facets(a_kind, L_kind, x.shape[1], cells, x.ravel(), cd, kd,
                                   fc.shape[0], fc, fl, _ptr(mask), float(alpha), _ptr(g), rp, cl,
                                   _ptr(vals), _ptr(b))
    if rc != 0:
        raise RuntimeError(f"oracle assemble_facets failed ({rc})")
    return vals, b


def assemble_system_facets(a_kind, L_kind, fa_kind, fL_kind, x, cells, cell_dofs, facets,
                           row_ptr, col, bc_mask, bc_value, coeff_dofs=None, a_c0=None,
                           L_c0=None, alpha=1.0, g=None, facet_mask=None, x0=None):
    """SystemAssembler with the facet tensors merged into each cell tensor before apply_bc;
    x0 as in assemble_system (SystemAssembler.cpp:252-275)."""
    if x0 is not None:
        m = np.asarray(bc_mask) != 0
        bc_value = np.where(m, np.asarray(x0, dtype=np.float64) - np.asarray(bc_value, dtype=np.float64), 0.0)
    x = _c(x, np.float64)
    cells = _c(cells, np.int32)
    cd = _c(cell_dofs, np.int32)
    kd = _c(cells if coeff_dofs is None else coeff_dofs, np.int32)
    fc, fl = (_c(a, np.int32) for a in facets)
    mask = None if facet_mask is None else _c(facet_mask, np.uint8)
    arrs = [None if a is None else _c(a, np.float64) for a in (a_c0, L_c0, g)]
    vals = np.zeros(col.shape[0])
    b = np.zeros(row_ptr.shape[0] - 1)
    rc = lib().orc_assemble_system_facets(a_kind, L_kind, fa_kind, fL_kind, x.shape[1],
                                          cells.shape[0], cells, x.ravel(), cd, kd,
                                          _ptr(arrs[0]), _ptr(arrs[1]), float(alpha),
                                          _ptr(arrs[2]), fc.shape[0], fc, fl, _ptr(mask),
                                          _c(bc_mask, np.uint8), _c(bc_value, np.float64),
                                          _c(row_ptr, np.int32), _c(col, np.int32), _ptr(vals),
                                          _ptr(b))
    if rc != 0:
        raise RuntimeError(f"oracle assemble_system_facets failed ({rc})")
    return vals, b


def adjoint_p1(x, cells, cell_dofs, kappa, f, u):
    x = _c(x, np.float64)
    cells = _c(cells, np.int32)
    cd = _c(cell_dofs, np.int32)
    nd = kappa.shape[0]
    dk = np.empty(nd)
    dx = np.empty_like(x)
    J = lib().orc_adjoint_p1(x.shape[1], cells.shape[0], cells, x.ravel(), x.shape[0], cd, nd,
                             _c(kappa, np.float64), _c(f, np.float64), _c(u, np.float64), dk,
                             dx.ravel())
    return J, dk, dx


def adjoint_vp2(x, cells, cell_dofs, coeff_dofs, mu, lmbda, f3, u):
    x = _c(x, np.float64)
    cells = _c(cells, np.int32)
    nk = mu.shape[0]
    dmu = np.empty(nk)
    dla = np.empty(nk)
    dx = np.empty_like(x)
    J = lib().orc_adjoint_vp2(cells.shape[0], cells, x.ravel(), x.shape[0],
                              _c(cell_dofs, np.int32), _c(coeff_dofs, np.int32), nk,
                              _c(mu, np.float64), _c(lmbda, np.float64), _c(f3, np.float64),
                              _c(u, np.float64), dmu, dla, dx.ravel())
    return J, dmu, dla, dx


def frobenius(values):
    """EigenMatrix::norm("frobenius") (EigenMatrix.cpp:77-105): sqrt of sum of squares."""
    return float(np.sqrt(np.sum(np.asarray(values) ** 2)))


def boundary_vertices(x):
    """Vertices on the boundary of the unit interval/square/cube (DirichletBC 'on_boundary'
    for P1, topological method)."""
    x = np.asarray(x)
    return np.any((np.abs(x) < 1e-14) | (np.abs(x - 1.0) < 1e-14), axis=1)


# ---- node reordering (pure Python: small graphs only) -----------------------------------
def node_graph(cell_nodes, num_nodes):
    """Sorted adjacency lists of the node graph: two nodes are adjacent when they share a cell
    (the graph DofMapBuilder::compute_node_reordering builds, DofMapBuilder.cpp:1530-1558)."""
    nbr = [set() for _ in range(num_nodes)]
    for c in np.asarray(cell_nodes):
        for a in c:
            for b in c:
                if a != b:
                    nbr[int(a)].add(int(b))
    return [sorted(s) for s in nbr]


def _bfs_levels(adj, s):
    """Eccentricity of s and the first smallest-degree vertex of its last BFS level (the
    rcm_queue bookkeeping Boost's pseudo_peripheral_pair reads)."""
    seen = {s}
    level = [s]
    ecc = -1
    last_min = s
    while level:
        last_min = level[0]
        nxt = []
        for u in level:
            if len(adj[u]) < len(adj[last_min]):
                last_min = u
            for v in adj[u]:
                if v not in seen:
                    seen.add(v)
                    nxt.append(v)
        ecc += 1
        level = nxt
    return ecc, last_min


MAX_QUEUE = 16


def _stable_by_degree(adj, seq):
    # Boost sorts each vertex's newly queued neighbours with std::sort by degree. For lists of
    # at most 16 libstdc++'s std::sort is an insertion sort, which keeps equal degrees in queue
    # order, as this stable sort does. Longer lists go through introsort, whose order of equal
    # degrees this oracle does not restate: it refuses them rather than guess (parity of the
    # reverse Cuthill-McKee numbering is unpinned there).
    if len(seq) > MAX_QUEUE:
        raise ValueError(f"a vertex queues {len(seq)} > {MAX_QUEUE} new neighbours at once: "
                         "Boost's order of equal degrees is not restated for such graphs")
    return sorted(seq, key=lambda v: len(adj[v]))


def cuthill_mckee(adj, start):
    """Cuthill-McKee order of start's component: breadth first, each vertex's unvisited
    neighbours (ascending index) queued in ascending degree."""
    order = [start]
    seen = {start}
    head = 0
    while head < len(order):
        u = order[head]
        head += 1
        new = [v for v in adj[u] if v not in seen]
        seen.update(new)
        order.extend(_stable_by_degree(adj, new))
    return order


def reverse_cuthill_mckee(adj):
    """old_to_new of the reverse Cuthill-McKee numbering as BoostGraphOrdering::
    compute_cuthill_mckee(graph, true) returns it (BoostGraphOrdering.cpp:34-75, with Boost's
    cuthill_mckee_ordering): components in order of their lowest vertex, each started at the
    pseudo-peripheral vertex of George and Liu (find_starting_node), the sequence reversed."""
    n = len(adj)
    if sum(len(a) for a in adj) == 0:
        return np.arange(n, dtype=np.int32)
    done = set()
    order = []
    for r0 in range(n):
        if r0 in done:
            continue
        r = r0
        ecc_r, x = _bfs_levels(adj, r)
        ecc_x, y = _bfs_levels(adj, x)
        while ecc_x > ecc_r:
            r, ecc_r, x = x, ecc_x, y
            ecc_x, y = _bfs_levels(adj, x)
        comp = cuthill_mckee(adj, x)
        done.update(comp)
        order.extend(comp)
    old_to_new = np.empty(n, dtype=np.int32)
    for i, v in enumerate(order):
        old_to_new[v] = n - 1 - i
    return old_to_new


def bandwidth(adj, old_to_new=None):
    """max |new(u) - new(v)| over the edges."""
    p = np.arange(len(adj)) if old_to_new is None else np.asarray(old_to_new)
    return max((abs(int(p[u]) - int(p[v])) for u in range(len(adj)) for v in adj[u]), default=0)


# ---- cell colouring (pure Python; small meshes) ---------------------------------------------
def cell_entity_graph(cells, dim):
    """GraphBuilder::local_graph(mesh, tdim, dim) (dolfin/graph/GraphBuilder.cpp:124-153) for a
    simplicial mesh: the entities of dimension dim of a cell are its (dim+1)-vertex subsets
    (as sorted tuples, unique per vertex set as in TopologyComputation), and a cell's
    neighbours are the other cells holding one of its entities. Returns a list of sets."""
    from itertools import combinations
    cells = np.asarray(cells)
    owners = {}
    for c, vs in enumerate(cells.tolist()):
        for e in combinations(sorted(vs), dim + 1):
            owners.setdefault(e, []).append(c)
    graph = [set() for _ in range(len(cells))]
    for c, vs in enumerate(cells.tolist()):
        for e in combinations(sorted(vs), dim + 1):
            graph[c].update(n for n in owners[e] if n != c)
    return graph


def sequential_vertex_coloring(graph):
    """Boost's sequential_vertex_coloring (boost/graph/sequential_vertex_coloring.hpp, Boost
    1.56+; called from dolfin/graph/BoostGraphColoring.h:112-113), restated from its published
    description: every vertex starts with the colour V - 1, vertices are visited in index
    order, the colours of all neighbours are marked, and the vertex takes the smallest
    unmarked colour below the current number of colours, or opens a new one. Returns the
    colours and their number."""
    V = len(graph)
    color = [V - 1] * V
    mark = [-1] * max(V, 1)
    max_color = 0
    for i in range(V):
        for u in graph[i]:
            mark[color[u]] = i
        k = 0
        while k < max_color and mark[k] == i:
            k += 1
        if k == max_color:
            max_color += 1
        color[i] = k
    return np.array(color, dtype=np.int64), max_color
