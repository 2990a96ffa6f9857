"""ctypes binding of libglu_hip.so (the C-ABI declared in include/glu_hip.h).

Loading fails loudly when the library is missing: there is no CPU
fallback for any entry point.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.environ.get("GLU_HIP_LIB", os.path.join(_PKG, "libglu_hip.so"))

GLU_OK = 0
STATUS_NAMES = {
    -1: "GLU_ERR_GENERAL", -2: "GLU_ERR_ARGUMENT", -3: "GLU_ERR_MEMORY", -4: "GLU_ERR_DEVICE",
    -5: "GLU_ERR_STRUCT_SINGULAR", -6: "GLU_ERR_NUMERIC", -7: "GLU_ERR_INVALID",
    -8: "GLU_ERR_DUPLICATE", -9: "GLU_ERR_EMPTY",
}

# glu_field enumerators
F_SYM_C_PTR, F_SYM_R_IDX, F_L_COL_PTR, F_CSR_R_PTR, F_CSR_C_IDX, F_CSR_DIAG_PTR, F_VAL, \
    F_LEVEL_PTR, F_LEVEL_IDX, F_ROW_PERM, F_ROW_PERM_INV, F_COL_PERM, F_COL_PERM_INV, \
    F_ROW_SCALE, F_COL_SCALE_PERM, F_DUMP_AX, F_DUMP_AI, F_DUMP_AP = range(18)

EXPORTED = [
    "glu_last_error", "glu_analyze", "glu_analyze_many", "glu_analysis_destroy", "glu_analysis_info", "glu_analysis_get",
    "glu_plan_create", "glu_plan_create_from_analysis", "glu_plan_destroy", "glu_plan_info",
    "glu_factor", "glu_factor_host", "glu_solve", "glu_refresh", "glu_refine", "glu_adjoint_values",
    "glu_batch_solve", "glu_batch_factor", "glu_abft_sums", "glu_abft_check",
    "glu_group_create", "glu_group_destroy", "glu_group_set_split", "glu_group_solve", "glu_refine_noalloc",
]


class GluError(RuntimeError):
    def __init__(self, status: int, where: str, text: str):
        name = STATUS_NAMES.get(status, str(status))
        super().__init__(f"{where}: {name}: {text}")
        self.status = status


class Pattern(C.Structure):
    _fields_ = [
        ("n", C.c_uint32), ("nnz", C.c_uint32), ("num_lev", C.c_int32),
        ("sym_c_ptr", C.c_void_p), ("sym_r_idx", C.c_void_p), ("l_col_ptr", C.c_void_p),
        ("csr_r_ptr", C.c_void_p), ("csr_c_idx", C.c_void_p), ("csr_diag_ptr", C.c_void_p),
        ("level_ptr", C.c_void_p), ("level_idx", C.c_void_p),
    ]


class PlanOpts(C.Structure):
    _fields_ = [("fuse_units", C.c_int32), ("block_threads", C.c_int32), ("use_graph", C.c_int32),
                ("lds_windows", C.c_int32), ("solve_dataflow", C.c_int32), ("krylov", C.c_int32),
                ("lds_window_values", C.c_int32), ("batch_grid_levels", C.c_int32)]


_lib = None


def lib():
    """Load libglu_hip.so (raises OSError if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OSError(f"libglu_hip.so not built at {LIB_PATH}; run __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    P, I, U, D = C.c_void_p, C.c_int, C.c_uint32, C.c_double
    st = C.c_int
    sig = {
        "glu_last_error": (C.c_char_p, []),
        "glu_analyze": (st, [U, U, P, P, P, C.POINTER(P)]),
        "glu_analyze_many": (st, [I, P, P, P, P, P, I, P, P]),
        "glu_analysis_destroy": (None, [P]),
        "glu_analysis_info": (st, [P, P]),
        "glu_analysis_get": (st, [P, I, P, C.c_size_t]),
        "glu_plan_create": (st, [C.POINTER(Pattern), C.POINTER(PlanOpts), C.POINTER(P)]),
        "glu_plan_create_from_analysis": (st, [P, C.POINTER(PlanOpts), C.POINTER(P)]),
        "glu_plan_destroy": (None, [P]),
        "glu_plan_info": (st, [P, P]),
        "glu_factor": (st, [P, P, I, P]),
        "glu_factor_host": (st, [P, P, I]),
        "glu_solve": (st, [P, P, P, P, I, P]),
        "glu_refresh": (st, [P, P, P, I, P]),
        "glu_refine": (st, [P, P, P, P, P, D, I, I, P, P, P]),
        "glu_refine_noalloc": (st, [P, P, P, P, P, D, I, I, P, P, P]),
        "glu_adjoint_values": (st, [P, P, P, P, P]),
        "glu_batch_solve": (st, [P, I, P, P, P, P, D, I, P, P, C.c_uint, P]),
        "glu_batch_factor": (st, [P, I, P, P, C.c_uint, P]),
        "glu_abft_sums": (st, [P, I, P, P, P]),
        "glu_abft_check": (st, [P, I, P, P, C.c_float, P, P, P]),
        "glu_group_create": (st, [P, I, C.POINTER(P)]),
        "glu_group_destroy": (None, [P]),
        "glu_group_set_split": (st, [P, I]),
        "glu_group_solve": (st, [P, P, P, P, P, D, I, P, P, C.c_uint, P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(status: int, where: str):
    if status != GLU_OK:
        raise GluError(status, where, lib().glu_last_error().decode())


def ptr(a) -> int:
    """Address of a numpy array's data or a torch tensor's storage."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    return a.data_ptr()
