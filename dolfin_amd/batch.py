"""Batched assembly of many independent forms in one graph-captured launch.

Each member is a bilinear form (P1 kappa-Poisson on a tetrahedral mesh, or vector-P2
elasticity) with its own function space, coefficients and CSR tensor. The library builds
unit tables over all members' row blocks and captures two kernels (vector-P2 members, then
P1 members, each at its own occupancy) in one hipGraph; assemble() replays it. grid_per_cu=0
launches one workgroup per row block, grid_per_cu>0 that many persistent workgroups per CU.
Members without a structure and plan yet get them first, several at a time (prepare_engines).
Coefficient vectors may be updated in place between calls.
"""
import ctypes
from concurrent.futures import ThreadPoolExecutor

import torch

from . import _lib
from .fem import engine_for
from .la import Matrix


def prepare_engines(forms, plan_defaults=(0, 0, 0), workers=8):
    """Builds every member's structure and plan (engine_for), `workers` members at a time: the
    host parts of the builds (vector-P2 rank tables and block plans) run concurrently, the
    device parts queue on the default stream. plan_defaults (P1 rows per block, P1 threads,
    vector-P2 values per block; 0 = library default) apply to the builds, in each worker."""
    L = _lib.load()

    def one(a):
        _lib.check(L.dfa_set_plan_defaults(*(int(v) for v in plan_defaults)))
        try:
            engine_for(a.V)
        finally:
            L.dfa_set_plan_defaults(0, 0, 0)

    pending = [a for a in forms if not a.V._engines.get(False)]
    if workers <= 1 or len(pending) <= 1:
        for a in pending:
            one(a)
        return
    torch.cuda.synchronize()
    with ThreadPoolExecutor(max_workers=workers) as ex:
        list(ex.map(one, pending))
    torch.cuda.synchronize()


class BatchAssembler:
    def __init__(self, forms, grid_per_cu=0, plan_defaults=(0, 0, 0), workers=8):
        if not forms:
            raise ValueError("empty batch")
        L = _lib.load()
        self.forms = list(forms)
        prepare_engines(self.forms, plan_defaults, workers)
        self.tensors = []
        engines, kinds, xs, c0s, c1s, vals = [], [], [], [], [], []
        self._keep = []
        for a in self.forms:
            if a.rank != 2 or a.kind not in (_lib.FORM_P1_POISSON, _lib.FORM_VP2_ELASTICITY):
                raise ValueError("batch members must be P1 Poisson or vector-P2 elasticity forms")
            eng = engine_for(a.V)
            A = Matrix()
            A.init(eng)
            self.tensors.append(A)
            x = a.V.mesh().coordinates_device()
            cv = a.coefficient_vectors()
            engines.append(eng.handle.value)
            kinds.append(a.kind)
            xs.append(x.data_ptr())
            c0s.append(cv[0].data_ptr())
            c1s.append(cv[1].data_ptr() if len(cv) > 1 else 0)
            vals.append(A.values.data_ptr())
            self._keep.append((eng, x, cv))
        n = len(self.forms)
        arr = lambda v: (ctypes.c_void_p * n)(*v)
        h = ctypes.c_void_p()
        _lib.check(L.dfa_batch_create(ctypes.byref(h), n, arr(engines), (ctypes.c_int32 * n)(*kinds),
                                      arr(xs), arr(c0s), arr(c1s), arr(vals), grid_per_cu,
                                      _lib.stream()))
        self._h = h
        u, g, l, c = ctypes.c_int64(), ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int64()
        _lib.check(L.dfa_batch_info(h, ctypes.byref(u), ctypes.byref(g), ctypes.byref(l),
                                    ctypes.byref(c)))
        self.info = {"units": u.value, "grid": g.value, "lds_bytes": l.value, "cells": c.value,
                     "members": n}

    def assemble(self):
        """Replay the captured launch on the current stream; returns the member tensors."""
        _lib.check(_lib.load().dfa_batch_launch(self._h, _lib.stream()))
        return self.tensors

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.load().dfa_batch_destroy(h)
            except Exception:
                pass
            self._h = None
