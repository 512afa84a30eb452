"""ctypes binding of libdolfin_amd.so (the C ABI in include/dolfin_amd.h).

torch is imported first on purpose: its bundled HIP runtime carries the same soname
(libamdhip64.so.7) as /opt/rocm's, so the library binds to the runtime torch already loaded
and device pointers/streams from torch are valid in it. There is no CPU fallback: if the
library is missing, every call raises.
"""
import ctypes
import os

import torch  # noqa: F401  (load torch's HIP runtime before the library)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("DFA_LIBRARY") or os.path.join(_HERE, "libdolfin_amd.so")

ELEMENT_P1 = 1
ELEMENT_VECTOR_P2 = 2

FORM_P1_POISSON = 0
FORM_P1_LOAD = 1
FORM_P1_MASS = 2
FORM_VP2_ELASTICITY = 3
FORM_VP2_BODY_FORCE = 4
FORM_P1_FACET_MASS = 7
FORM_P1_FACET_LOAD = 8
FORM_VP2_FACET_TRACTION = 9

SCATTER_AUTO = 0
SCATTER_ATOMIC = 1
SCATTER_BLOCKED = 2
SCATTER_COLORED = 3

ADD_VALUES = 1

EXPORTS = ["dfa_last_error", "dfa_version", "dfa_assembler_create", "dfa_assembler_destroy",
           "dfa_csr_dims", "dfa_csr_structure", "dfa_assemble_matrix", "dfa_assemble_vector",
           "dfa_assemble_system", "dfa_adjoint", "dfa_plan_info", "dfa_set_plan_defaults",
           "dfa_batch_create", "dfa_batch_launch", "dfa_batch_info", "dfa_batch_destroy",
           "dfa_exterior_facets", "dfa_assemble_facets", "dfa_csr_spmv", "dfa_csr_diagonal", "dfa_cg_solve",
           "dfa_node_reordering", "dfa_graph_reordering", "dfa_mesh_coloring", "dfa_color_info",
           "dfa_device_info", "dfa_assemble_system_x0", "dfa_dirichlet_x0", "dfa_pass_kernel", "dfa_plan_builder",
           "dfa_csr_add_local", "dfa_vector_add_local", "dfa_assemble_system_x0_noalloc"]

_lib = None

_vp = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64


def load():
    """Load and declare the library (no GPU work happens here)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run `python -m dolfin_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    L.dfa_last_error.restype = ctypes.c_char_p
    L.dfa_version.restype = ctypes.c_char_p
    L.dfa_assembler_create.argtypes = [ctypes.POINTER(_vp), _i32, _i64, _i64, _vp, _i32, _i32, _vp,
                                       _i64, _vp, _i64, _vp, _i32, _vp]
    L.dfa_assembler_destroy.argtypes = [_vp]
    L.dfa_csr_dims.argtypes = [_vp, ctypes.POINTER(_i64), ctypes.POINTER(_i64)]
    L.dfa_csr_structure.argtypes = [_vp, _vp, _vp, _vp]
    L.dfa_assemble_matrix.argtypes = [_vp, _i32, _vp, _vp, _vp, _vp, _i32, _i32, _vp]
    L.dfa_assemble_vector.argtypes = [_vp, _i32, _vp, _vp, _vp, _vp, _i32, _i32, _vp]
    L.dfa_assemble_system.argtypes = [_vp, _i32, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                      _i32, _i32, _vp]
    L.dfa_assemble_system_x0.argtypes = [_vp, _i32, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                         _vp, _vp, _i32, _i32, _vp]
    L.dfa_assemble_system_x0_noalloc.argtypes = L.dfa_assemble_system_x0.argtypes
    L.dfa_dirichlet_x0.argtypes = [_i64, _vp, _vp, _vp, _vp, _vp]
    L.dfa_csr_add_local.argtypes = [_vp, _vp, _i32, _vp, _i32, _vp, _vp, _vp]
    L.dfa_vector_add_local.argtypes = [_vp, _i64, _i32, _vp, _vp, _vp]
    L.dfa_pass_kernel.argtypes = [_vp, _i32, _i32, _i32, _i32, ctypes.c_char_p, _i32]
    L.dfa_plan_builder.argtypes = [_vp, ctypes.c_void_p]
    L.dfa_adjoint.argtypes = [_vp, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]
    L.dfa_plan_info.argtypes = [_vp, _i32, ctypes.POINTER(_i64), ctypes.POINTER(_i64),
                                ctypes.POINTER(_i64), ctypes.POINTER(_i32)]
    L.dfa_set_plan_defaults.argtypes = [_i32, _i32, _i32]
    _pp = ctypes.POINTER(ctypes.c_void_p)
    L.dfa_batch_create.argtypes = [ctypes.POINTER(_vp), _i32, _pp, ctypes.POINTER(_i32), _pp, _pp, _pp,
                                   _pp, _i32, _vp]
    L.dfa_batch_launch.argtypes = [_vp, _vp]
    L.dfa_batch_info.argtypes = [_vp, ctypes.POINTER(_i64), ctypes.POINTER(_i32), ctypes.POINTER(_i32),
                                 ctypes.POINTER(_i64)]
    L.dfa_batch_destroy.argtypes = [_vp]
    L.dfa_exterior_facets.argtypes = [_vp, ctypes.POINTER(_i64), _vp, _vp]
    L.dfa_csr_spmv.argtypes = [_vp, _vp, ctypes.c_double, _vp, ctypes.c_double, _vp, _vp]
    L.dfa_csr_diagonal.argtypes = [_vp, _vp, _vp, _vp]
    L.dfa_cg_solve.argtypes = [_vp, _vp, _vp, _vp, _i32, ctypes.c_double, ctypes.c_double, _i64, _i32,
                               ctypes.POINTER(_i64), _vp]
    L.dfa_node_reordering.argtypes = [_i64, _i32, _vp, _i64, _vp]
    L.dfa_graph_reordering.argtypes = [_i64, _vp, _vp, _vp]
    L.dfa_mesh_coloring.argtypes = [_i64, _i32, _vp, _i64, _i32, _vp, ctypes.POINTER(_i32)]
    L.dfa_color_info.argtypes = [_vp, ctypes.POINTER(_i32), _vp, _vp]
    L.dfa_device_info.argtypes = [_i32, _vp]
    L.dfa_assemble_facets.argtypes = [_vp, _i32, _i32, _vp, ctypes.c_double, _vp, _vp, _vp, _vp,
                                      _vp, _vp, _vp]
    for name in EXPORTS:
        if name not in ("dfa_last_error", "dfa_version"):
            getattr(L, name).restype = ctypes.c_int
    _lib = L
    return L


def check(rc):
    if rc != 0:
        msg = load().dfa_last_error().decode()
        raise RuntimeError(f"dolfin_amd error {rc}: {msg}")


def ptr(t):
    """Raw device/host pointer of a torch tensor or numpy array (None -> NULL)."""
    if t is None:
        return None
    if isinstance(t, torch.Tensor):
        return ctypes.c_void_p(t.data_ptr())
    return t.ctypes.data_as(ctypes.c_void_p)


def stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
