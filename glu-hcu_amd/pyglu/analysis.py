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
        self._adopt(h)

    @classmethod
    def _from_handle(cls, h):
        obj = cls.__new__(cls)
        obj._adopt(h)
        return obj

    def _adopt(self, h):
        L = capi.lib()
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


def analyze_many(mats, threads: int = 0):
    """Analyse independent CSC matrices ``[(data, indices, indptr, n), ...]``
    on host threads (``glu_analyze_many``; 0 = one per core).  Returns one
    :class:`Analysis` per matrix, each identical to ``Analysis(*mat)``."""
    L = capi.lib()
    cnt = len(mats)
    if cnt == 0:
        return []
    keep = []
    ns = np.zeros(cnt, dtype=np.uint32)
    nnzs = np.zeros(cnt, dtype=np.uint32)
    pax = (C.c_void_p * cnt)()
    pai = (C.c_void_p * cnt)()
    pap = (C.c_void_p * cnt)()
    for k, (data, indices, indptr, n) in enumerate(mats):
        ax = np.ascontiguousarray(data, dtype=np.float64)
        ai = np.ascontiguousarray(indices, dtype=np.uint32)
        ap = np.ascontiguousarray(indptr, dtype=np.uint32)
        if ap.size != n + 1:
            raise ValueError(f"matrix {k}: indptr length {ap.size} inconsistent with n={n}")
        if ai.size != ax.size:
            raise ValueError(f"matrix {k}: data and indices length mismatch")
        keep += [ax, ai, ap]
        ns[k], nnzs[k] = n, ax.size
        pax[k], pai[k], pap[k] = capi.ptr(ax), capi.ptr(ai), capi.ptr(ap)
    out = (C.c_void_p * cnt)()
    status = np.zeros(cnt, dtype=np.int32)
    rc = L.glu_analyze_many(cnt, capi.ptr(ns), capi.ptr(nnzs), pax, pai, pap, int(threads), out,
                            capi.ptr(status))
    if rc != capi.GLU_OK:
        for k in range(cnt):
            if out[k]:
                L.glu_analysis_destroy(C.c_void_p(out[k]))
        capi.check(rc, "glu_analyze_many")
    return [Analysis._from_handle(C.c_void_p(out[k])) for k in range(cnt)]
