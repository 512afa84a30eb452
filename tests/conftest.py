import ctypes
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


def _gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def pytest_collection_modifyitems(config, items):
    if _gpu_available():
        return
    skip = pytest.mark.skip(reason="no GPU in this environment")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)


@pytest.fixture(scope="session")
def oracle():
    import oracle as O
    O.build()
    return O


_PLAN = None


@pytest.fixture(scope="session")
def plan_lib():
    """tests/plan_check.cpp + the library's host structure code, built with g++."""
    global _PLAN
    if _PLAN is None:
        out = os.path.join(ROOT, "tests", "_build")
        os.makedirs(out, exist_ok=True)
        so = os.path.join(out, "libplan_check.so")
        srcs = [os.path.join(ROOT, "tests", "plan_check.cpp"),
                os.path.join(ROOT, "dolfin_amd", "csrc", "structure.cpp"),
                os.path.join(ROOT, "oracle", "oracle.cpp")]
        if not os.path.exists(so) or any(os.path.getmtime(s) > os.path.getmtime(so) for s in srcs):
            subprocess.check_call(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-pthread",
                                   "-o", so] + srcs)
        L = ctypes.CDLL(so)
        vp = ctypes.c_void_p
        L.plan_check_p1.argtypes = [ctypes.c_int32, ctypes.c_int64, ctypes.c_int64, vp, vp, vp, vp,
                                    ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, vp, vp,
                                    ctypes.c_int64, vp]
        L.plan_check_p1.restype = ctypes.c_int
        L.plan_check_p1_chains.argtypes = [ctypes.c_int64, ctypes.c_int64, vp, vp, vp, vp, vp,
                                           ctypes.c_int64, ctypes.c_int32, vp, vp, ctypes.c_int64,
                                           vp]
        L.plan_check_p1_chains.restype = ctypes.c_int
        L.facets_host.argtypes = [ctypes.c_int32, ctypes.c_int64, vp, vp, vp]
        L.facets_host.restype = ctypes.c_int64
        L.spmv_blocks_host.argtypes = [vp, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, vp]
        L.spmv_blocks_host.restype = ctypes.c_int64
        L.plan_check_vp2.argtypes = [ctypes.c_int64, vp, vp, vp, vp, ctypes.c_int64, vp, vp,
                                     ctypes.c_int32, vp, ctypes.c_int64, vp]
        L.plan_check_vp2.restype = ctypes.c_int
        L.csr_host.argtypes = [ctypes.c_int64, ctypes.c_int32, vp, ctypes.c_int64, ctypes.c_int32,
                               vp, vp, ctypes.c_int64]
        L.csr_host.restype = ctypes.c_int64
        _PLAN = L
    return _PLAN


def p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def host_csr(plan_lib, cell_dofs, ndofs, keep_diagonal=False):
    cd = np.ascontiguousarray(cell_dofs, dtype=np.int32)
    cap = int(cd.shape[0]) * cd.shape[1] * cd.shape[1] + ndofs
    rp = np.empty(ndofs + 1, np.int32)
    col = np.empty(cap, np.int32)
    nnz = plan_lib.csr_host(cd.shape[0], cd.shape[1], p(cd), ndofs, int(keep_diagonal), p(rp),
                            p(col), cap)
    assert nnz >= 0
    return rp, col[:nnz].copy()
