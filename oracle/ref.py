"""ctypes access to oracle/_ref/libglu_ref.so (test infrastructure only).

The library is the reference's own host code (NICSLU analysis, my_DumpA,
Symbolic_Matrix) compiled from the reference sources by oracle/Makefile.
It is used to check the product's host pipeline bit-for-bit and to time the
NICSLU CPU baseline.  When the library is absent, ``available()`` is False
and callers skip.
"""

from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_ref", "libglu_ref.so")
_lib = None


def available() -> bool:
    return os.path.exists(_LIB_PATH)


def _load():
    global _lib
    if _lib is None:
        lib = C.CDLL(_LIB_PATH)
        P = C.c_void_p
        lib.ref_analyze.argtypes = [C.c_uint, C.c_uint, P, P, P, C.POINTER(C.c_void_p)]
        lib.ref_analyze.restype = C.c_int
        lib.ref_free.argtypes = [P]
        lib.ref_sizes.argtypes = [P, P]
        lib.ref_perms.argtypes = [P] * 7
        lib.ref_dump.argtypes = [P] * 4
        lib.ref_symbolic.argtypes = [P] * 10
        lib.ref_solve.argtypes = [P] * 4
        lib.ref_abft.argtypes = [P] * 4
        lib.ref_abft.restype = C.c_int
        lib.ref_refact_values.argtypes = [P] * 4
        lib.ref_refact_values.restype = C.c_int
        lib.ref_nicslu_time.argtypes = [C.c_uint, C.c_uint, P, P, P, P, C.c_int, C.c_int, P, P]
        lib.ref_nicslu_time.restype = C.c_int
        _lib = lib
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


@dataclass
class RefAnalysis:
    n: int
    nnz_a: int
    nnz_lu: int
    num_lev: int
    row_perm: np.ndarray
    row_perm_inv: np.ndarray
    col_perm: np.ndarray
    col_perm_inv: np.ndarray
    row_scale: np.ndarray
    col_scale_perm: np.ndarray
    dump_ax: np.ndarray
    dump_ai: np.ndarray
    dump_ap: np.ndarray
    sym_c_ptr: np.ndarray
    sym_r_idx: np.ndarray
    l_col_ptr: np.ndarray
    csr_r_ptr: np.ndarray
    csr_c_idx: np.ndarray
    csr_diag_ptr: np.ndarray
    val: np.ndarray
    level_ptr: np.ndarray
    level_idx: np.ndarray
    _handle: int = 0

    def solve(self, lu: np.ndarray, rhs: np.ndarray) -> np.ndarray:
        """Reference FP32 host solve (src/symbolic.cc:200-251)."""
        lu = np.ascontiguousarray(lu, dtype=np.float32)
        rhs = np.ascontiguousarray(rhs, dtype=np.float32)
        x = np.empty(self.n, dtype=np.float32)
        _load().ref_solve(C.c_void_p(self._handle), _p(lu), _p(rhs), _p(x))
        return x

    def abft(self, val0: np.ndarray, lu: np.ndarray):
        """The reference's ABFT (src/symbolic.cc:163-197): CCA of val0, then
        the check of the factors lu.  Returns (cca, first failing column or
        -1 for a pass)."""
        val0 = np.ascontiguousarray(val0, dtype=np.float32)
        lu = np.ascontiguousarray(lu, dtype=np.float32)
        cca = np.empty(self.n, dtype=np.float32)
        first = _load().ref_abft(C.c_void_p(self._handle), _p(val0), _p(lu), _p(cca))
        return cca, int(first)

    def refact_values(self, ax0: np.ndarray):
        """The reference's values for a refactorization with new values ax0
        (caller CSC order): NICSLU's _I_NicsLU_MC64ScaleForRefact +
        _I_NicsLU_Scale, my_DumpA, then predictLU.  Returns (dumped FP64
        values, FP32 values in the LU pattern)."""
        ax0 = np.ascontiguousarray(ax0, dtype=np.float64)
        assert ax0.size == self.nnz_a
        dump = np.empty(self.nnz_a)
        val = np.empty(self.nnz_lu, dtype=np.float32)
        rc = _load().ref_refact_values(C.c_void_p(self._handle), _p(ax0), _p(dump), _p(val))
        if rc != 0:
            raise RefError(rc)
        return dump, val

    def close(self):
        if self._handle:
            _load().ref_free(C.c_void_p(self._handle))
            self._handle = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class RefError(RuntimeError):
    def __init__(self, code):
        super().__init__(f"reference analysis failed with code {code}")
        self.code = code


def analyze(data, indices, indptr, n) -> RefAnalysis:
    lib = _load()
    ax = np.ascontiguousarray(data, dtype=np.float64)
    ai = np.ascontiguousarray(indices, dtype=np.uint32)
    ap = np.ascontiguousarray(indptr, dtype=np.uint32)
    h = C.c_void_p()
    rc = lib.ref_analyze(n, ax.size, _p(ax), _p(ai), _p(ap), C.byref(h))
    if rc != 0:
        raise RefError(rc)
    sizes = np.zeros(4, dtype=np.int64)
    lib.ref_sizes(h, _p(sizes))
    n_, nnz_a, nnz_lu, nl = (int(s) for s in sizes)
    u = lambda k: np.empty(k, dtype=np.uint32)
    rp, rpi, cp, cpi = u(n_), u(n_), u(n_), u(n_)
    rs, csp = np.empty(n_), np.empty(n_)
    lib.ref_perms(h, _p(rp), _p(rpi), _p(cp), _p(cpi), _p(rs), _p(csp))
    dax, dai, dap = np.empty(nnz_a), u(nnz_a), u(n_ + 1)
    lib.ref_dump(h, _p(dax), _p(dai), _p(dap))
    scp, sri, lcp = u(n_ + 1), u(nnz_lu), u(n_)
    crp, cci, cdp = u(n_ + 1), u(nnz_lu), u(n_)
    val = np.empty(nnz_lu, dtype=np.float32)
    lp = np.empty(nl + 1, dtype=np.int32)
    li = np.empty(n_, dtype=np.int32)
    lib.ref_symbolic(h, _p(scp), _p(sri), _p(lcp), _p(crp), _p(cci), _p(cdp), _p(val),
                     _p(lp), _p(li))
    return RefAnalysis(n_, nnz_a, nnz_lu, nl, rp, rpi, cp, cpi, rs, csp, dax, dai, dap,
                       scp, sri, lcp, crp, cci, cdp, val, lp, li, h.value)


def nicslu_time(data, indices, indptr, n, value_sets, threads=1):
    """Time NICSLU Analyze/Factorize/ReFactorize(_MT)/Solve on the host."""
    lib = _load()
    ax = np.ascontiguousarray(data, dtype=np.float64)
    ai = np.ascontiguousarray(indices, dtype=np.uint32)
    ap = np.ascontiguousarray(indptr, dtype=np.uint32)
    vals = np.ascontiguousarray(value_sets, dtype=np.float64).reshape(-1, ax.size)
    times = np.zeros(4)
    x = np.empty(n)
    rc = lib.ref_nicslu_time(n, ax.size, _p(ax), _p(ai), _p(ap), _p(vals), vals.shape[0],
                             threads, _p(times), _p(x))
    if rc != 0:
        raise RefError(rc)
    return {"analyze_s": times[0], "factorize_s": times[1], "refactor_s": times[2],
            "solve_s": times[3], "x": x}
