"""Device-resident tensors: the GenericMatrix / GenericVector subset the assembler uses.

Matrix mirrors EigenMatrix (row-major CSR, sorted columns, int32 indices; dolfin/la/
EigenMatrix.h:46-230): init from the engine's structure, zero, apply (no-op: values live in
HBM), nnz, size, norm, data() export (EigenMatrix.cpp:341-355), mult (y = A x,
EigenMatrix.cpp:230-255) and the diagonal. solve() is dolfin.solve(A, x, b, "cg", "jacobi")
(python/dolfin/la/solver.py:27, dolfin/la/solve.cpp:40) on the device: Jacobi-preconditioned
conjugate gradients as EigenKrylovSolver runs them, with every vector in HBM.
"""

import ctypes

import numpy as np
import torch

from . import _lib


class Matrix:
    def __init__(self):
        self._engine = None
        self.values = None
        self._row_ptr = None
        self._col = None

    # GenericTensor::empty (AssemblerBase.cpp:150 decides first-call initialisation on it)
    def empty(self):
        return self.values is None

    def init(self, engine):
        self._engine = engine
        self.values = torch.zeros(engine.nnz, dtype=torch.float64, device=engine.device)
        self._row_ptr = None
        self._col = None

    def engine(self):
        return self._engine

    def size(self, dim=0):
        return 0 if self._engine is None else self._engine.num_rows

    def nnz(self):
        return 0 if self._engine is None else self._engine.nnz

    def local_range(self, dim=0):
        """GenericTensor::local_range (GenericTensor.h:69): one process owns every row."""
        return (0, self.size(dim))

    def zero(self):
        if self.values is not None:
            self.values.zero_()

    def add_local(self, block, rows, cols):
        """GenericMatrix::add_local(block, rows) from the host (GenericMatrix.h:106-112;
        EigenMatrix::add, EigenMatrix.cpp:156-166): A[rows[i], cols[j]] += block[i, j]. Every
        entry must be in the sparsity pattern."""
        r = np.ascontiguousarray(rows, dtype=np.int32)
        c = np.ascontiguousarray(cols, dtype=np.int32)
        blk = np.ascontiguousarray(block, dtype=np.float64).reshape(r.shape[0], c.shape[0])
        _lib.check(_lib.load().dfa_csr_add_local(self._engine.handle, _lib.ptr(self.values),
                                                 r.shape[0], _lib.ptr(r), c.shape[0], _lib.ptr(c),
                                                 _lib.ptr(blk), _lib.stream()))

    def apply(self, mode="add"):
        return None

    def structure(self):
        """(row_ptr, col) as device int32 tensors."""
        if self._row_ptr is None:
            self._row_ptr, self._col = self._engine.structure()
        return self._row_ptr, self._col

    def norm(self, norm_type="frobenius"):
        if norm_type != "frobenius":
            raise ValueError("only the Frobenius norm is provided")
        return float(torch.sqrt(torch.sum(self.values * self.values)).item())

    def data(self):
        """(row_ptr, col, values) as numpy arrays (copy to host)."""
        rp, col = self.structure()
        torch.cuda.synchronize()
        return rp.cpu().numpy(), col.cpu().numpy(), self.values.cpu().numpy()

    def mult(self, x, y=None, alpha=1.0, beta=0.0):
        """y = alpha * A x + beta * y (device tensors or Vectors); returns y."""
        xv = x.values if isinstance(x, Vector) else x
        if y is None:
            y = torch.empty(self.size(), dtype=torch.float64, device=self.values.device)
            beta = 0.0
        yv = y.values if isinstance(y, Vector) else y
        _lib.check(_lib.load().dfa_csr_spmv(self._engine.handle, _lib.ptr(self.values),
                                            float(alpha), _lib.ptr(xv), float(beta), _lib.ptr(yv),
                                            _lib.stream()))
        return y

    def diagonal(self):
        d = torch.empty(self.size(), dtype=torch.float64, device=self.values.device)
        _lib.check(_lib.load().dfa_csr_diagonal(self._engine.handle, _lib.ptr(self.values),
                                                _lib.ptr(d), _lib.stream()))
        return d

    def array(self):
        rp, col, v = self.data()
        n = rp.shape[0] - 1
        A = np.zeros((n, n))
        rows = np.repeat(np.arange(n), np.diff(rp))
        A[rows, col] = v
        return A


class Vector:
    def __init__(self, n=None, device="cuda"):
        self.values = None if n is None else torch.zeros(n, dtype=torch.float64, device=device)

    def empty(self):
        return self.values is None

    def init(self, n, device="cuda"):
        self.values = torch.zeros(n, dtype=torch.float64, device=device)

    def size(self):
        return 0 if self.values is None else self.values.shape[0]

    def local_range(self):
        return (0, self.size())

    def zero(self):
        if self.values is not None:
            self.values.zero_()

    def add_local(self, block, rows):
        """GenericVector::add_local (EigenVector.cpp:160-165): b[rows[i]] += block[i]."""
        r = np.ascontiguousarray(rows, dtype=np.int32)
        blk = np.ascontiguousarray(block, dtype=np.float64).reshape(r.shape[0])
        _lib.check(_lib.load().dfa_vector_add_local(_lib.ptr(self.values), self.size(), r.shape[0],
                                                    _lib.ptr(r), _lib.ptr(blk), _lib.stream()))

    def apply(self, mode="add"):
        return None

    def norm(self, norm_type="l2"):
        if norm_type == "l2":
            return float(torch.sqrt(torch.sum(self.values * self.values)).item())
        if norm_type == "linf":
            return float(torch.max(torch.abs(self.values)).item())
        if norm_type == "l1":
            return float(torch.sum(torch.abs(self.values)).item())
        raise ValueError(norm_type)

    def get_local(self):
        return self.values.cpu().numpy()


def solve(A, x, b, method="cg", preconditioner="jacobi", rtol=1e-10, atol=0.0, maxiter=None,
          check_every=10):
    """Solve A x = b on the device (x holds the initial guess and receives the solution).
    method "cg" with preconditioner "jacobi" or "none" (EigenKrylovSolver's methods, dolfin/la/
    EigenKrylovSolver.cpp:43-56); stops at the first iteration with |r| <= max(rtol |b|, atol).
    Returns that iteration count (dolfin::solve returns the same count); raises when the method
    does not converge or breaks down (p.Ap = 0 with r != 0).

    The iteration runs in the library (dfa_cg_solve: the SpMV and three fused vector kernels
    per iteration, deterministic reductions); the residual is tested on the device every
    iteration, the host reads the result every check_every iterations, and the iterate is
    frozen from the converged iteration on, so the answer and the count do not depend on
    check_every."""
    if method != "cg":
        raise ValueError("only the conjugate gradient method is provided")
    if preconditioner not in ("jacobi", "none", "default"):
        raise ValueError("preconditioner must be 'jacobi' or 'none'")
    xv = x.values if isinstance(x, Vector) else x
    bv = b.values if isinstance(b, Vector) else b
    n = A.size()
    for t, name in ((xv, "x"), (bv, "b")):
        if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous() and t.numel() == n):
            raise ValueError(f"{name} must be a contiguous float64 device vector of size {n}")
    its = ctypes.c_int64(0)
    rc = _lib.load().dfa_cg_solve(A._engine.handle, _lib.ptr(A.values), _lib.ptr(xv), _lib.ptr(bv),
                                  int(preconditioner != "none"), float(rtol), float(atol),
                                  -1 if maxiter is None else int(maxiter), int(check_every),
                                  ctypes.byref(its), _lib.stream())
    if rc != 0:
        raise RuntimeError(_lib.load().dfa_last_error().decode())
    return int(its.value)
