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

    def pass_kernel(self, a_kind, L_kind=-1, has_bc=False, scatter=_lib.SCATTER_AUTO):
        """Name and template arguments of the kernel an assembly with these forms would run
        (a_kind or L_kind -1: matrix or vector alone)."""
        buf = ctypes.create_string_buffer(256)
        _lib.check(_lib.load().dfa_pass_kernel(self._h, a_kind, L_kind, int(bool(has_bc)), scatter,
                                               buf, 256))
        return buf.value.decode()

    def plan_builder(self):
        """"device" when the block plan (P1, or vector P2) was built by the device planner,
        "host" otherwise; raises when a DFA_PLAN_VERIFY=1 creation found the two builds
        different."""
        v = ctypes.c_int32()
        _lib.check(_lib.load().dfa_plan_builder(self._h, ctypes.byref(v)))
        return "device" if v.value else "host"

    def color_ptr(self):
        """Offsets of each colour's cells in the coloured scatter (P1; built on first use)."""
        L = _lib.load()
        n = ctypes.c_int32()
        _lib.check(L.dfa_color_info(self._h, ctypes.byref(n), None, _lib.stream()))
        ptr = np.empty(n.value + 1, dtype=np.int64)
        _lib.check(L.dfa_color_info(self._h, ctypes.byref(n), _lib.ptr(ptr), _lib.stream()))
        return ptr


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


# ---- exterior-facet forms -----------------------------------------------------------------
def exterior_facets(V):
    """(facet_cell, facet_local) of the exterior facets of V's mesh in the library's order
    (sorted by cell, then by local facet = the facet opposite that local vertex)."""
    eng = engine_for(V)
    lib = _lib.load()
    n = ctypes.c_int64()
    _lib.check(lib.dfa_exterior_facets(eng.handle, ctypes.byref(n), None, None))
    fc = np.empty(n.value, np.int32)
    fl = np.empty(n.value, np.int32)
    _lib.check(lib.dfa_exterior_facets(eng.handle, ctypes.byref(n), _lib.ptr(fc), _lib.ptr(fl)))
    return fc, fl


def mark_exterior_facets(V, inside, check_midpoint=True):
    """SubDomain::mark over the exterior facets (dolfin/mesh/SubDomain.cpp:182-282): a facet is
    marked when inside(x) holds at every vertex and, with check_midpoint, at its midpoint.
    inside takes an array of points [n, gdim] and returns booleans [n]. Returns a uint8 mask
    in exterior_facets(V) order."""
    mesh = V.mesh()
    fc, fl = exterior_facets(V)
    cells = mesh.cells()
    x = mesh.coordinates()
    nvc = cells.shape[1]
    cols = np.array([[a for a in range(nvc) if a != f] for f in range(nvc)])
    fv = cells[fc[:, None], cols[fl]]               # [nf, gdim] facet vertices
    pts = x[fv]                                       # [nf, gdim, gdim]
    ok = np.ones(fc.shape[0], dtype=bool)
    for j in range(fv.shape[1]):
        ok &= np.asarray(inside(pts[:, j, :]), dtype=bool)
    if check_midpoint:
        ok &= np.asarray(inside(pts.mean(axis=1)), dtype=bool)
    return ok.astype(np.uint8)


class FacetForm:
    """An exterior-facet form, optionally over the marked facets only (an exterior-facet
    subdomain): P1 alpha*u*v*ds (rank 2) or g*v*ds (rank 1); vector-P2 inner(t, v)*ds with a
    constant traction t (rank 1)."""

    def __init__(self, kind, rank, V, alpha=0.0, g=None, marker=None, traction=None):
        if kind == _lib.FORM_VP2_FACET_TRACTION:
            _vp2(V)
            t3 = np.asarray(traction, dtype=np.float64).reshape(3)
            self.traction = torch.from_numpy(t3.copy()).cuda()
        else:
            _p1(V)
            self.traction = None
        self.kind, self.rank, self.V = kind, rank, V
        self.alpha = float(alpha)
        self.g = g
        if g is not None and (not isinstance(g, Function) or g.function_space().element != _lib.ELEMENT_P1
                              or g.function_space().mesh() is not V.mesh()):
            raise ValueError("g must be a P1 Function on the same mesh")
        self._marker = marker
        self._mask = None

    def facet_mask(self):
        """Device uint8 mask over the exterior facets, or None for all of them."""
        if self._marker is None:
            return None
        if self._mask is None:
            m = self._marker
            if callable(m):
                m = mark_exterior_facets(self.V, m)
            m = np.ascontiguousarray(np.asarray(m, dtype=np.uint8))
            self._mask = torch.from_numpy(m).cuda()
        return self._mask

    def __add__(self, other):
        return FormSum([self]) + other


class FormSum:
    """Sum of cell and exterior-facet forms of one rank on one space (a ufl Form with dx and
    ds integrals)."""

    def __init__(self, parts):
        self.parts = list(parts)
        self.rank = self.parts[0].rank
        self.V = self.parts[0].V
        for f in self.parts:
            if f.rank != self.rank or f.V is not self.V:
                raise ValueError("summed forms need the same rank and function space")

    def __add__(self, other):
        return FormSum(self.parts + (other.parts if isinstance(other, FormSum) else [other]))

    def cell_parts(self):
        return [f for f in self.parts if isinstance(f, Form)]

    def facet_parts(self):
        return [f for f in self.parts if isinstance(f, FacetForm)]


Form.__add__ = lambda self, other: FormSum([self]) + other


def facet_mass_form(V, alpha=1.0, marker=None):
    """alpha*u*v*ds (alpha constant), over the exterior facets selected by marker."""
    return FacetForm(_lib.FORM_P1_FACET_MASS, 2, V, alpha=alpha, marker=marker)


def facet_load_form(V, g, marker=None):
    """g*v*ds (g a P1 Function), over the exterior facets selected by marker."""
    return FacetForm(_lib.FORM_P1_FACET_LOAD, 1, V, g=g, marker=marker)


def facet_traction_form(V, t3, marker=None):
    """inner(t, v)*ds on a vector-P2 space, t a constant 3-vector (the Neumann term of
    elasticity), over the exterior facets selected by marker."""
    return FacetForm(_lib.FORM_VP2_FACET_TRACTION, 1, V, traction=t3, marker=marker)


def _split(form):
    if isinstance(form, FormSum):
        return form.cell_parts(), form.facet_parts()
    if isinstance(form, FacetForm):
        return [], [form]
    return [form], []


def _assemble_facets(eng, x, a_facets, L_facets, A_values, b_values, mask=None, vals=None):
    """Adds exterior-facet terms; a facet matrix and vector with different markers run as
    separate passes."""
    lib = _lib.load()
    groups = []
    for fa in a_facets:
        groups.append((fa, None))
    for fl in L_facets:
        groups.append((None, fl))
    # one pass when the matrix and vector terms share their facet selection
    if len(a_facets) == 1 and len(L_facets) == 1 and a_facets[0]._marker is L_facets[0]._marker:
        groups = [(a_facets[0], L_facets[0])]
    for fa, fl in groups:
        fmask = (fa or fl).facet_mask()
        _lib.check(lib.dfa_assemble_facets(
            eng.handle, fa.kind if fa is not None else -1, fl.kind if fl is not None else -1,
            _lib.ptr(x), fa.alpha if fa is not None else 0.0,
            _lib.ptr(None if fl is None else (fl.traction if fl.traction is not None else fl.g.vector())),
            _lib.ptr(fmask), _lib.ptr(mask),
            _lib.ptr(vals), _lib.ptr(A_values if fa is not None else None),
            _lib.ptr(b_values), _lib.stream()))


# ---- boundary conditions ------------------------------------------------------------------
def boundary_facet_nodes(V):
    """Dofs on boundary facets (DirichletBC 'on_boundary', topological method): the facets of
    exactly one cell, taken from the library's exterior-facet list (cell, local facet k =
    the facet opposite local vertex k); P1 -> their vertices' dofs, vector P2 -> the dofs of
    their vertices and edges, every component."""
    fc, fl = exterior_facets(V)
    fc = fc.astype(np.int64)
    cd = V.cell_dofs()
    nvc = V.mesh().cells().shape[1]
    # local vertices of each local facet: on_facet[k, a] = (a != k)
    on_facet = ~np.eye(nvc, dtype=bool)
    if V.element == _lib.ELEMENT_P1:
        vsel = on_facet[fl]                                      # [nf, nvc]
        return np.unique(cd[fc][vsel])
    # vector P2, component-major local dofs comp*10 + node; nodes 0-3 vertices, 4-9 the UFC
    # edges (2,3),(1,3),(1,2),(0,3),(0,2),(0,1): an edge lies on facet k when k is not one of
    # its vertices
    loc = np.array([[2, 3], [1, 3], [1, 2], [0, 3], [0, 2], [0, 1]])
    edge_on = np.array([[k not in e for e in loc] for k in range(4)])   # [4, 6]
    nodes_on = np.concatenate([on_facet, edge_on], axis=1)[fl]        # [nf, 10]
    c = cd[fc]                                                          # [nf, 30]
    sel = np.concatenate([nodes_on] * 3, axis=1)
    return np.unique(c[sel])


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
        cells, facets = _split(a)
        if facets:
            # cell integrals first (zeroing unless add_values), then the exterior facets
            # (Assembler.cpp:52-103 order)
            if len(cells) > 1:
                raise ValueError("at most one cell-integral form per sum")
            eng = engine_for(a.V, self.keep_diagonal)
            fin, self.finalize_tensor = self.finalize_tensor, False
            try:
                if cells:
                    self.assemble(A, cells[0])
                else:
                    if a.rank == 2 and (A.empty() or A.engine() is not eng):
                        A.init(eng)
                    if a.rank == 1 and (A.empty() or A.size() != a.V.dim()):
                        A.init(a.V.dim())
                    if not self.add_values:
                        A.zero()
            finally:
                self.finalize_tensor = fin
            x = a.V.mesh().coordinates_device()
            if a.rank == 2:
                _assemble_facets(eng, x, facets, [], A.values, None)
            else:
                _assemble_facets(eng, x, [], facets, None, A.values)
            if self.finalize_tensor:
                A.apply("add")
            return A
        a = cells[0]
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
        a_cells, self.a_facets = _split(a)
        L_cells, self.L_facets = _split(L)
        if len(a_cells) != 1 or len(L_cells) > 1:
            raise ValueError("SystemAssembler needs one cell-integral bilinear form and at most "
                             "one cell-integral linear form (plus exterior-facet forms)")
        if L_cells:
            L = L_cells[0]
        else:
            # a linear form of facet terms only: its cell part is f*v*dx with f = 0
            if a.V.element == _lib.ELEMENT_P1:
                L = load_form(a.V, Function(a.V))
            else:
                L = body_force_form(a.V, (0.0, 0.0, 0.0))
        self.a, self.L = a_cells[0], L
        self.bcs = [] if bcs is None else (list(bcs) if isinstance(bcs, (list, tuple)) else [bcs])
        self.scatter = scatter
        # AssemblerBase flags (dolfin/fem/AssemblerBase.h:46-67), honoured as the reference's
        # SystemAssembler does through init_global_tensor
        self.add_values = False
        self.finalize_tensor = True
        self.keep_diagonal = False
        self._x0_vals = None

    def assemble(self, A=None, b=None, x0=None):
        """assemble(A, b) / assemble(A) / assemble(b=b), and with x0 the incremental variants
        assemble(A, b, x0) / assemble(b=b, x0=x0) (SystemAssembler.h:63-84): every Dirichlet
        value g becomes x0[dof] - g (SystemAssembler.cpp:252-275)."""
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
        flags = _lib.ADD_VALUES if self.add_values else 0
        has_facets = bool(self.a_facets or self.L_facets)
        x0v = None
        if x0 is not None:
            x0v = x0.values if isinstance(x0, Vector) else x0
            if x0v.shape[0] != V.dim():
                raise ValueError("x0 must have one entry per dof")
            if mask is not None and has_facets:
                # the facet pass needs the same shifted values: form them once
                if self._x0_vals is None or self._x0_vals.shape[0] != V.dim():
                    self._x0_vals = torch.empty(V.dim(), dtype=torch.float64, device="cuda")
                _lib.check(lib.dfa_dirichlet_x0(V.dim(), _lib.ptr(mask), _lib.ptr(vals),
                                                _lib.ptr(x0v), _lib.ptr(self._x0_vals),
                                                _lib.stream()))
                vals, x0v = self._x0_vals, None
        args = (eng.handle, self.a.kind, self.L.kind, _lib.ptr(x),
                _lib.ptr(ac[0] if len(ac) > 0 else None), _lib.ptr(ac[1] if len(ac) > 1 else None),
                _lib.ptr(lc[0] if lc else None), _lib.ptr(self.L.f3), _lib.ptr(mask),
                _lib.ptr(vals))
        outs = (_lib.ptr(None if A is None else A.values), _lib.ptr(None if b is None else b.values),
                self.scatter, flags, _lib.stream())
        if x0v is not None:
            _lib.check(lib.dfa_assemble_system_x0(*args, _lib.ptr(x0v), *outs))
        else:
            _lib.check(lib.dfa_assemble_system(*args, *outs))
        if has_facets:
            # facet tensors belong to their cell's tensor before apply_bc
            # (SystemAssembler.cpp:421-481): Dirichlet rows keep the cell pass's values and
            # Dirichlet columns are lifted into b
            if len(self.a_facets) > 1 or len(self.L_facets) > 1:
                raise ValueError("at most one exterior-facet form each in a and L")
            _assemble_facets(eng, x, self.a_facets, self.L_facets,
                             None if A is None else A.values, None if b is None else b.values,
                             mask=mask, vals=vals)
        if self.finalize_tensor:
            for T in (A, b):
                if T is not None:
                    T.apply("add")
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


def assemble_system(a, L, bcs=None, x0=None, add_values=False, finalize_tensor=True,
                    keep_diagonal=False, A_tensor=None, b_tensor=None, scatter=_lib.SCATTER_AUTO):
    """python/dolfin/fem/assembling.py:322-385 (same argument order, without the form
    compiler parameters and backend): flags set on the SystemAssembler, x0 selects the
    incremental assemble(A, b, x0)."""
    A = Matrix() if A_tensor is None else A_tensor
    b = Vector() if b_tensor is None else b_tensor
    asm = SystemAssembler(a, L, bcs, scatter)
    asm.add_values = add_values
    asm.finalize_tensor = finalize_tensor
    asm.keep_diagonal = keep_diagonal
    asm.assemble(A, b, x0=x0)
    return A, b
