"""Host analysis (MC64 + AMD + symbolic) through the C-ABI.

``Analysis`` owns a ``glu_analysis*`` and exposes every array of the
reference's host pipeline: the SNicsLU permutations and scales
(src/nicslu/source/nicslu.h:44-70), the my_DumpA matrix
(src/preprocess/preprocess.c:8-100) and the Symbolic_Matrix fields
(include/symbolic.h:12-23).
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import _capi as capi

_FIELDS = {
    # name: (enum, dtype, length key)
    "sym_c_ptr": (capi.F_SYM_C_PTR, np.uint32, "n+1"),
    "sym_r_idx": (capi.F_SYM_R_IDX, np.uint32, "nnz_lu"),
    "l_col_ptr": (capi.F_L_COL_PTR, np.uint32, "n"),
    "csr_r_ptr": (capi.F_CSR_R_PTR, np.uint32, "n+1"),
    "csr_c_idx": (capi.F_CSR_C_IDX, np.uint32, "nnz_lu"),
    "csr_diag_ptr": (capi.F_CSR_DIAG_PTR, np.uint32, "n"),
    "val": (capi.F_VAL, np.float32, "nnz_lu"),
    "level_ptr": (capi.F_LEVEL_PTR, np.int32, "num_lev+1"),
    "level_idx": (capi.F_LEVEL_IDX, np.int32, "n"),
    "row_perm": (capi.F_ROW_PERM, np.uint32, "n"),
    "row_perm_inv": (capi.F_ROW_PERM_INV, np.uint32, "n"),
    "col_perm": (capi.F_COL_PERM, np.uint32, "n"),
    "col_perm_inv": (capi.F_COL_PERM_INV, np.uint32, "n"),
    "row_scale": (capi.F_ROW_SCALE, np.float64, "n"),
    "col_scale_perm": (capi.F_COL_SCALE_PERM, np.float64, "n"),
    "dump_ax": (capi.F_DUMP_AX, np.float64, "nnz_a"),
    "dump_ai": (capi.F_DUMP_AI, np.uint32, "nnz_a"),
    "dump_ap": (capi.F_DUMP_AP, np.uint32, "n+1"),
}


class Analysis:
    """Result of ``glu_analyze`` on one CSC matrix (copied)."""

    def __init__(self, data, indices, indptr, n: int):
        L = capi.lib()
        ax = np.ascontiguousarray(data, dtype=np.float64)
        ai = np.ascontiguousarray(indices, dtype=np.uint32)
        ap = np.ascontiguousarray(indptr, dtype=np.uint32)
        if ap.size != n + 1:
            raise ValueError(f"indptr length {ap.size} inconsistent with n={n}")
        if ai.size != ax.size:
            raise ValueError("data and indices length mismatch")
        h = C.c_void_p()
        capi.check(L.glu_analyze(n, ax.size, capi.ptr(ax), capi.ptr(ai), capi.ptr(ap), C.byref(h)),
                   "glu_analyze")
        self._h = h
        info = np.zeros(8, dtype=np.int64)
        capi.check(L.glu_analysis_info(h, capi.ptr(info)), "glu_analysis_info")
        self.n = int(info[0])
        self.nnz_a = int(info[1])
        self.nnz_lu = int(info[2])
        self.num_lev = int(info[3])
        self.updates = int(info[4])
        self._cache = {}

    @property
    def handle(self):
        return self._h

    def _len(self, key):
        return {"n": self.n, "n+1": self.n + 1, "nnz_lu": self.nnz_lu, "nnz_a": self.nnz_a,
                "num_lev+1": self.num_lev + 1}[key]

    def __getattr__(self, name):
        if name.startswith("_") or name not in _FIELDS:
            raise AttributeError(name)
        if name not in self._cache:
            enum, dt, key = _FIELDS[name]
            out = np.empty(self._len(key), dtype=dt)
            capi.check(capi.lib().glu_analysis_get(self._h, enum, capi.ptr(out), out.nbytes),
                       f"glu_analysis_get({name})")
            self._cache[name] = out
        return self._cache[name]

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            capi.lib().glu_analysis_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
