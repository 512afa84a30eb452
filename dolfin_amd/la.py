"""Device-resident tensors: the GenericMatrix / GenericVector subset the assembler uses.

Matrix mirrors EigenMatrix (row-major CSR, sorted columns, int32 indices; dolfin/la/
EigenMatrix.h:46-230): init from the engine's structure, zero, apply (no-op: values live in
HBM), nnz, size, norm, data() export (EigenMatrix.cpp:341-355).
"""
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

    def zero(self):
        if self.values is not None:
            self.values.zero_()

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

    def zero(self):
        if self.values is not None:
            self.values.zero_()

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
