"""The allocation-free entry point dfa_assemble_system_x0_noalloc against the allocating
dfa_assemble_system_x0, through the C ABI, on two calls in a row (the second on a smaller
mesh), plus its refusals (too many dofs for its fixed device array; the coloured scatter before
its colour plan exists). Bar: values within relative Frobenius / L2 error 1e-12 (FP64; the
blocked kernels' LDS accumulation order may differ between passes)."""
import numpy as np
import pytest
import torch

import dolfin_amd as dfa
from dolfin_amd import _lib, workloads

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _rel(a, b):
    return (torch.linalg.norm(a - b) / torch.linalg.norm(b)).item()


def _p1_args(n, seed):
    mesh, V, kappa, f = workloads.p1_poisson_3d(n, seed=seed)
    rng = np.random.default_rng(seed)
    bc = dfa.DirichletBC(V, 0.0)
    g = np.zeros(V.dim())
    g[bc.dofs] = rng.uniform(-1, 1, bc.dofs.shape[0])
    eng = dfa.engine_for(V)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    return dict(eng=eng, V=V, a=_lib.FORM_P1_POISSON, L=_lib.FORM_P1_LOAD,
                x=mesh.coordinates_device(), c0=dev(kappa), c1=None, Lc0=dev(f), f3=None,
                mask=bc.mask, g=dev(g), x0=dev(rng.uniform(-1, 1, V.dim())))


def _vp2_args(n, seed):
    mesh, V, mu, lmbda, f3 = workloads.vp2_elasticity_3d(n, seed=seed)
    rng = np.random.default_rng(seed)
    bc = dfa.DirichletBC(V, 0.0)
    g = np.zeros(V.dim())
    g[bc.dofs] = rng.uniform(-1, 1, bc.dofs.shape[0])
    eng = dfa.engine_for(V)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    return dict(eng=eng, V=V, a=_lib.FORM_VP2_ELASTICITY, L=_lib.FORM_VP2_BODY_FORCE,
                x=mesh.coordinates_device(), c0=dev(mu), c1=dev(lmbda), Lc0=None,
                f3=np.ascontiguousarray(f3, dtype=np.float64), mask=bc.mask, g=dev(g),
                x0=dev(rng.uniform(-1, 1, V.dim())))


def _call(fn, c, scatter=_lib.SCATTER_AUTO):
    A = torch.full((c["eng"].nnz,), 3.0, dtype=torch.float64, device="cuda")  # overwritten
    b = torch.full((c["V"].dim(),), 3.0, dtype=torch.float64, device="cuda")
    rc = fn(c["eng"].handle, c["a"], c["L"], _lib.ptr(c["x"]), _lib.ptr(c["c0"]),
            _lib.ptr(c["c1"]), _lib.ptr(c["Lc0"]), _lib.ptr(c["f3"]), _lib.ptr(c["mask"]),
            _lib.ptr(c["g"]), _lib.ptr(c["x0"]), _lib.ptr(A), _lib.ptr(b), scatter, 0,
            _lib.stream())
    torch.cuda.synchronize()
    return rc, A, b


@pytest.mark.parametrize("kind", ["p1", "vp2"])
def test_noalloc_matches_allocating_entry_point(kind):
    lib = _lib.load()
    make = _p1_args if kind == "p1" else _vp2_args
    # the second call has a smaller input: the shared device array holds the first call's
    # (longer) contents, which the second call must not read
    for n, seed in ((7, 1), (3, 2)) if kind == "p1" else ((3, 1), (2, 2)):
        print(f"case noalloc {kind} n={n}", flush=True)
        c = make(n, seed)
        rc0, A0, b0 = _call(lib.dfa_assemble_system_x0, c)
        rc1, A1, b1 = _call(lib.dfa_assemble_system_x0_noalloc, c)
        assert rc0 == 0 and rc1 == 0, lib.dfa_last_error().decode()
        assert _rel(A1, A0) < TOL and _rel(b1, b0) < TOL


def test_noalloc_refusals():
    lib = _lib.load()
    c = _p1_args(3, 3)
    # the coloured scatter before its colour plan exists
    rc, _, _ = _call(lib.dfa_assemble_system_x0_noalloc, c, scatter=_lib.SCATTER_COLORED)
    assert rc == -4 and "colour plan" in lib.dfa_last_error().decode()
    rc0, A0, b0 = _call(lib.dfa_assemble_system_x0, c, scatter=_lib.SCATTER_COLORED)
    rc1, A1, b1 = _call(lib.dfa_assemble_system_x0_noalloc, c, scatter=_lib.SCATTER_COLORED)
    assert rc0 == 0 and rc1 == 0
    assert _rel(A1, A0) < TOL and _rel(b1, b0) < TOL

    # more dofs than the fixed array holds: a tiny mesh whose space carries rows in no cell
    class Padded:
        element = _lib.ELEMENT_P1

        def __init__(self, mesh, extra):
            self._mesh, self._extra = mesh, extra
            self._W = dfa.FunctionSpace(mesh, "P", 1)

        def mesh(self):
            return self._mesh

        def cell_dofs(self):
            return self._mesh.cells()

        def dim(self):
            return self._mesh.num_vertices() + self._extra

        def coefficient_space(self):
            return self._W

    mesh = dfa.UnitCubeMesh(2, 2, 2)
    V = Padded(mesh, 4194304)
    eng = dfa.fem.Engine(V)
    big = dict(c, eng=eng, V=V, x=mesh.coordinates_device(),
               c0=torch.ones(mesh.num_vertices(), dtype=torch.float64, device="cuda"),
               Lc0=torch.ones(mesh.num_vertices(), dtype=torch.float64, device="cuda"),
               mask=torch.zeros(V.dim(), dtype=torch.uint8, device="cuda"),
               g=torch.zeros(V.dim(), dtype=torch.float64, device="cuda"),
               x0=torch.zeros(V.dim(), dtype=torch.float64, device="cuda"))
    rc, _, _ = _call(lib.dfa_assemble_system_x0_noalloc, big)
    assert rc == -3 and "exceed" in lib.dfa_last_error().decode()
    rc, _, _ = _call(lib.dfa_assemble_system_x0, big)
    assert rc == 0
