"""Assemblers: the Python mirror of dolfin.fem (Assembler, SystemAssembler, assemble,
assemble_system, DirichletBC) on top of the C ABI.

Semantics follow the reference:
  assemble(form, tensor=None)           python/dolfin/fem/assembling.py:79-220
  Assembler flags add_values / finalize_tensor / keep_diagonal
                                        dolfin/fem/AssemblerBase.h:46-67
  first call builds the structure, later calls only zero + reassemble
                                        dolfin/fem/AssemblerBase.cpp:137-285
  SystemAssembler(a, L, bcs).assemble(A, b) with per-cell symmetric Dirichlet lifting
                                        dolfin/fem/SystemAssembler.cpp:138-492,1036-1112
Unsupported forms raise (there is no fallback path).
"""
import ctypes

import numpy as np
import torch

from . import _lib
from .function import Function, FunctionSpace
from .la import Matrix, Vector


class Engine:
    """One dfa_assembler handle: (mesh, function space, keep_diagonal) -> device CSR structure
    and kernel plans, built once."""

    def __init__(self, V, keep_diagonal=False, device="cuda"):
        L = _lib.load()
        mesh = V.mesh()
        self.V = V
        self.device = device
        self.gdim = mesh.gdim()
        W = V.coefficient_space()
        h = ctypes.c_void_p()
        cells = np.ascontiguousarray(mesh.cells(), dtype=np.int32)
        cd = np.ascontiguousarray(V.cell_dofs(), dtype=np.int32)
        kd = np.ascontiguousarray(W.cell_dofs(), dtype=np.int32)
        x = np.ascontiguousarray(mesh.coordinates(), dtype=np.float64)
        torch.cuda.init()
        _lib.check(L.dfa_assembler_create(ctypes.byref(h), self.gdim, mesh.num_vertices(),
                                          mesh.num_cells(), _lib.ptr(cells), V.element,
                                          cd.shape[1], _lib.ptr(cd), V.dim(), _lib.ptr(kd),
                                          W.dim(), _lib.ptr(x), int(bool(keep_diagonal)),
                                          _lib.stream()))
        self._h = h
        nr, nnz = ctypes.c_int64(), ctypes.c_int64()
        _lib.check(L.dfa_csr_dims(h, ctypes.byref(nr), ctypes.byref(nnz)))
        self.num_rows = nr.value
        self.nnz = nnz.value

    @property
    def handle(self):
        return self._h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.load().dfa_assembler_destroy(h)
            except Exception:
                pass
            self._h = None

    def structure(self):
        rp = torch.empty(self.num_rows + 1, dtype=torch.int32, device=self.device)
        col = torch.empty(max(self.nnz, 1), dtype=torch.int32, device=self.device)
        _lib.check(_lib.load().dfa_csr_structure(self._h, _lib.ptr(rp), _lib.ptr(col), _lib.stream()))
        return rp, col[: self.nnz]

    def plan_info(self, scatter=_lib.SCATTER_AUTO):
        pb, nb, ca, lds = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int32()
        _lib.check(_lib.load().dfa_plan_info(self._h, scatter, ctypes.byref(pb), ctypes.byref(nb),
                                             ctypes.byref(ca), ctypes.byref(lds)))
        return {"plan_bytes": pb.value, "num_blocks": nb.value, "cell_appearances": ca.value,
                "lds_bytes": lds.value}


def engine_for(V, keep_diagonal=False):
    key = bool(keep_diagonal)
    e = V._engines.get(key)
    if e is None:
        e = Engine(V, keep_diagonal)
        V._engines[key] = e
    return e


# ---- forms -------------------------------------------------------------------------------
class Form:
    """A registered cell-integral form (the counterpart of an FFC-generated ufc::form)."""

    def __init__(self, kind, rank, V, coefficients=(), f3=None):
        self.kind = kind
        self.rank = rank
        self.V = V
        self.coefficients = tuple(coefficients)
        self.f3 = None if f3 is None else np.ascontiguousarray(f3, dtype=np.float64)
        for c in self.coefficients:
            if not isinstance(c, Function):
                raise TypeError("coefficients must be Functions")
            if c.function_space().element != _lib.ELEMENT_P1 or \
                    c.function_space().mesh() is not V.mesh():
                raise ValueError("coefficients must be P1 Functions on the same mesh")

    def function_space(self, i=0):
        return self.V

    def coefficient_vectors(self):
        return [c.vector() for c in self.coefficients]


def _p1(V):
    if V.element != _lib.ELEMENT_P1:
        raise ValueError("form needs a P1 space")


def _vp2(V):
    if V.element != _lib.ELEMENT_VECTOR_P2:
        raise ValueError("form needs a vector P2 space")


def poisson_form(V, kappa):
    """inner(kappa*grad(u), grad(v))*dx"""
    _p1(V)
    return Form(_lib.FORM_P1_POISSON, 2, V, (kappa,))


def mass_form(V):
    """u*v*dx"""
    _p1(V)
    return Form(_lib.FORM_P1_MASS, 2, V)


def load_form(V, f):
    """f*v*dx"""
    _p1(V)
    return Form(_lib.FORM_P1_LOAD, 1, V, (f,))


def elasticity_form(V, mu, lmbda):
    """inner(2*mu*eps(u) + lmbda*tr(eps(u))*I, eps(v))*dx"""
    _vp2(V)
    return Form(_lib.FORM_VP2_ELASTICITY, 2, V, (mu, lmbda))


def body_force_form(V, f3):
    """inner(f, v)*dx with a constant vector f"""
    _vp2(V)
    return Form(_lib.FORM_VP2_BODY_FORCE, 1, V, (), f3=f3)


# ---- boundary conditions ------------------------------------------------------------------
def boundary_facet_nodes(V):
    """Dofs on boundary facets (DirichletBC 'on_boundary', topological method): facets of
    exactly one cell; P1 -> their vertices, vector P2 -> their vertices and edges."""
    mesh = V.mesh()
    c = mesh.cells().astype(np.int64)
    n = c.shape[1]
    nv = mesh.num_vertices()
    facets = []
    for skip in range(n):
        f = np.sort(np.delete(c, skip, axis=1), axis=1)
        facets.append(f)
    F = np.concatenate(facets, axis=0)
    key = np.zeros(F.shape[0], dtype=np.int64)
    for k in range(F.shape[1]):
        key = key * nv + F[:, k]
    uniq, inv, cnt = np.unique(key, return_inverse=True, return_counts=True)
    bf = F[cnt[inv] == 1]
    onv = np.zeros(nv, dtype=bool)
    onv[bf.ravel()] = True
    if V.element == _lib.ELEMENT_P1:
        dofs_of_vertex = np.full(nv, -1, dtype=np.int64)
        dofs_of_vertex[mesh.cells().ravel()] = V.cell_dofs().ravel()
        return np.unique(dofs_of_vertex[onv & (dofs_of_vertex >= 0)])
    # vector P2: boundary edges are edges of boundary facets
    ce, ne = mesh.cell_edges()
    loc = np.array([[2, 3], [1, 3], [1, 2], [0, 3], [0, 2], [0, 1]])
    a = np.minimum(c[:, loc[:, 0]], c[:, loc[:, 1]])
    b = np.maximum(c[:, loc[:, 0]], c[:, loc[:, 1]])
    ekey = a * nv + b
    fe = set()
    for i, j in ((0, 1), (0, 2), (1, 2)):
        fe.update((bf[:, i] * nv + bf[:, j]).tolist())
    fe = np.fromiter(fe, dtype=np.int64)
    onedge = np.isin(ekey, fe)
    cd = V.cell_dofs()
    mask = np.zeros(V.dim(), dtype=bool)
    vert_on = onv[c]  # [nc, 4]
    for comp in range(3):
        mask[cd[:, comp * 10: comp * 10 + 4][vert_on]] = True
        mask[cd[:, comp * 10 + 4: comp * 10 + 10][onedge]] = True
    return np.nonzero(mask)[0]


class DirichletBC:
    """Dirichlet condition u = value on a set of dofs (default: all boundary-facet dofs)."""

    def __init__(self, V, value, dofs="on_boundary"):
        self.V = V
        if isinstance(dofs, str):
            if dofs != "on_boundary":
                raise ValueError("only 'on_boundary' or an explicit dof array is supported")
            dofs = boundary_facet_nodes(V)
        self.dofs = np.asarray(dofs, dtype=np.int64)
        vals = np.broadcast_to(np.asarray(value, dtype=np.float64), self.dofs.shape)
        mask = np.zeros(V.dim(), dtype=np.uint8)
        gv = np.zeros(V.dim(), dtype=np.float64)
        mask[self.dofs] = 1
        gv[self.dofs] = vals
        self.mask = torch.from_numpy(mask).cuda()
        self.values = torch.from_numpy(gv).cuda()


def _merge_bcs(V, bcs):
    if not bcs:
        return None, None
    if len(bcs) == 1:
        return bcs[0].mask, bcs[0].values
    mask = torch.zeros(V.dim(), dtype=torch.uint8, device="cuda")
    vals = torch.zeros(V.dim(), dtype=torch.float64, device="cuda")
    for bc in bcs:  # later conditions win, like the std::map insertion order in DOLFIN
        sel = bc.mask.bool()
        mask[sel] = 1
        vals[sel] = bc.values[sel]
    return mask, vals


# ---- assemblers --------------------------------------------------------------------------
class Assembler:
    """dolfin::Assembler (dolfin/fem/Assembler.h:66) for cell-integral forms."""

    def __init__(self, scatter=_lib.SCATTER_AUTO):
        self.add_values = False
        self.finalize_tensor = True
        self.keep_diagonal = False
        self.scatter = scatter

    def assemble(self, A, a):
        L = _lib.load()
        eng = engine_for(a.V, self.keep_diagonal)
        x = a.V.mesh().coordinates_device()
        flags = _lib.ADD_VALUES if self.add_values else 0
        cv = a.coefficient_vectors()
        if a.rank == 2:
            if A.empty() or A.engine() is not eng:
                A.init(eng)
            c0 = cv[0] if len(cv) > 0 else None
            c1 = cv[1] if len(cv) > 1 else None
            _lib.check(L.dfa_assemble_matrix(eng.handle, a.kind, _lib.ptr(x), _lib.ptr(c0),
                                             _lib.ptr(c1), _lib.ptr(A.values), self.scatter, flags,
                                             _lib.stream()))
        elif a.rank == 1:
            if A.empty() or A.size() != a.V.dim():
                A.init(a.V.dim())
            c0 = cv[0] if cv else None
            _lib.check(L.dfa_assemble_vector(eng.handle, a.kind, _lib.ptr(x), _lib.ptr(c0),
                                             _lib.ptr(a.f3), _lib.ptr(A.values), self.scatter,
                                             flags, _lib.stream()))
        else:
            raise ValueError("only rank-1 and rank-2 cell forms are supported")
        if self.finalize_tensor:
            A.apply("add")
        return A


class SystemAssembler:
    """dolfin::SystemAssembler (dolfin/fem/SystemAssembler.h:58) for one square bilinear form
    and one linear form with Dirichlet conditions applied per cell."""

    def __init__(self, a, L, bcs=None, scatter=_lib.SCATTER_AUTO):
        if a.rank != 2 or L.rank != 1:
            raise ValueError("SystemAssembler needs a bilinear and a linear form")
        if a.V is not L.V:
            raise ValueError("a and L must share the function space")
        self.a, self.L = a, L
        self.bcs = [] if bcs is None else (list(bcs) if isinstance(bcs, (list, tuple)) else [bcs])
        self.scatter = scatter
        self.keep_diagonal = False

    def assemble(self, A=None, b=None):
        if A is None and b is None:
            raise ValueError("nothing to assemble")
        lib = _lib.load()
        V = self.a.V
        eng = engine_for(V, self.keep_diagonal)
        if A is not None and (A.empty() or A.engine() is not eng):
            A.init(eng)
        if b is not None and (b.empty() or b.size() != V.dim()):
            b.init(V.dim())
        x = V.mesh().coordinates_device()
        ac = self.a.coefficient_vectors()
        lc = self.L.coefficient_vectors()
        mask, vals = _merge_bcs(V, self.bcs)
        _lib.check(lib.dfa_assemble_system(
            eng.handle, self.a.kind, self.L.kind, _lib.ptr(x),
            _lib.ptr(ac[0] if len(ac) > 0 else None), _lib.ptr(ac[1] if len(ac) > 1 else None),
            _lib.ptr(lc[0] if lc else None), _lib.ptr(self.L.f3), _lib.ptr(mask), _lib.ptr(vals),
            _lib.ptr(None if A is None else A.values), _lib.ptr(None if b is None else b.values),
            self.scatter, 0, _lib.stream()))
        return A, b


def assemble(form, tensor=None, add_values=False, finalize_tensor=True, keep_diagonal=False,
             scatter=_lib.SCATTER_AUTO):
    """python/dolfin/fem/assembling.py:79-220 for cell-integral forms: a fresh tensor unless
    one is passed; re-assembly into a passed tensor zeroes it first unless add_values."""
    if tensor is None:
        tensor = Matrix() if form.rank == 2 else Vector()
    asm = Assembler(scatter)
    asm.add_values = add_values
    asm.finalize_tensor = finalize_tensor
    asm.keep_diagonal = keep_diagonal
    return asm.assemble(tensor, form)


def assemble_system(a, L, bcs=None, A_tensor=None, b_tensor=None, scatter=_lib.SCATTER_AUTO):
    """python/dolfin/fem/assembling.py:322-385."""
    A = Matrix() if A_tensor is None else A_tensor
    b = Vector() if b_tensor is None else b_tensor
    SystemAssembler(a, L, bcs, scatter).assemble(A, b)
    return A, b
