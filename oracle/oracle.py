"""Python access to the oracle — TEST INFRASTRUCTURE ONLY.

Wraps oracle/libglu_oracle.so (the C restatement in glu_oracle.c) and adds
small dense references (numpy) used to pin it.  Only tests/,
__graft_entry__.smoke() and bench.py's CPU-baseline leg import this module.

Pinning status of each function (see tests/test_oracle.py):
  factor_f32    pinned: closed forms (diagonal, triangular, 2x2, tridiagonal
                recurrence), dense FP64 Doolittle on the same permuted
                matrix (n <= 300), FP64 oracle agreement at FP32 accuracy,
                the ABFT column-sum identity of src/symbolic.cc:172-197, the
                GESP quirks read from src/numeric.cu:133-146.
  factor_f32_sparse
                pinned: bit-exact against factor_f32 (+-GESP) on every
                analysed test matrix and on random patterns.
  solve_f32     pinned: bit-exact against the reference's own
                Symbolic_Matrix::solve compiled into oracle/_ref.
  factor_f64    pinned: dense Doolittle, scipy splu (natural order, no
                pivoting) on the dumped matrix.
  solve_f64     pinned: scipy.sparse.linalg.spsolve on the original A
                (A and A').
  abft_cca /    pinned: bit-exact CCA and the same first failing column as
  abft_check    the reference's own ABFTCalculateCCA / ABFTCheckResult
                (src/symbolic.cc:163-197, oracle/_ref built with GLU_DEBUG),
                on passing factors and on planted errors; closed forms
                (diagonal and unit-triangular matrices).
  refresh       pinned: bit-exact (FP64 dump and FP32 values) against the
                reference's own _I_NicsLU_MC64ScaleForRefact + my_DumpA +
                predictLU on drifted value sets; order 1 (the analysis
                product) must differ from it somewhere, so the check tells
                the two orders apart.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libglu_oracle.so")
_lib = None


class Pattern(C.Structure):
    _fields_ = [
        ("n", C.c_uint32), ("nnz", C.c_uint32), ("num_lev", C.c_int32),
        ("sym_c_ptr", C.c_void_p), ("sym_r_idx", C.c_void_p), ("l_col_ptr", C.c_void_p),
        ("csr_r_ptr", C.c_void_p), ("csr_c_idx", C.c_void_p), ("csr_diag_ptr", C.c_void_p),
        ("level_ptr", C.c_void_p), ("level_idx", C.c_void_p),
    ]


def available() -> bool:
    return os.path.exists(_LIB)


def _load():
    global _lib
    if _lib is None:
        if not available():
            raise OSError(f"{_LIB} not built (make -C oracle oracle)")
        L = C.CDLL(_LIB)
        P = C.c_void_p
        L.oracle_factor_f32.argtypes = [C.POINTER(Pattern), P, C.c_int]
        L.oracle_factor_f32.restype = C.c_float
        L.oracle_factor_f32_sparse.argtypes = [C.POINTER(Pattern), P, C.c_int]
        L.oracle_factor_f32_sparse.restype = C.c_float
        L.oracle_conflicts.argtypes = [C.POINTER(Pattern), P]
        L.oracle_conflicts.restype = C.c_uint32
        L.oracle_solve_f32.argtypes = [C.POINTER(Pattern), P, P, P, P, P, P, P]
        L.oracle_factor_f64.argtypes = [C.POINTER(Pattern), P]
        L.oracle_solve_f64.argtypes = [C.POINTER(Pattern)] + [P] * 9 + [C.c_int]
        L.oracle_abft_cca.argtypes = [C.POINTER(Pattern), P, P]
        L.oracle_abft_check.argtypes = [C.POINTER(Pattern), P, P, C.c_float, P, P]
        L.oracle_abft_check.restype = C.c_int32
        L.oracle_refresh.argtypes = [C.POINTER(Pattern), C.c_uint32] + [P] * 7 + [C.c_int, P, P]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class SymPattern:
    """Holds contiguous copies of the Symbolic_Matrix arrays of any object
    exposing them (a pyglu Analysis or an oracle.ref RefAnalysis)."""

    def __init__(self, src):
        u = lambda a: np.ascontiguousarray(a, dtype=np.uint32)
        i = lambda a: np.ascontiguousarray(a, dtype=np.int32)
        self.sym_c_ptr = u(src.sym_c_ptr)
        self.sym_r_idx = u(src.sym_r_idx)
        self.l_col_ptr = u(src.l_col_ptr)
        self.csr_r_ptr = u(src.csr_r_ptr)
        self.csr_c_idx = u(src.csr_c_idx)
        self.csr_diag_ptr = u(src.csr_diag_ptr)
        self.level_ptr = i(src.level_ptr)
        self.level_idx = i(src.level_idx)
        self.n = self.l_col_ptr.size
        self.nnz = self.sym_r_idx.size
        self.num_lev = self.level_ptr.size - 1
        self._c = Pattern(self.n, self.nnz, self.num_lev, *(
            a.ctypes.data for a in (self.sym_c_ptr, self.sym_r_idx, self.l_col_ptr, self.csr_r_ptr,
                                    self.csr_c_idx, self.csr_diag_ptr, self.level_ptr, self.level_idx)))

    @property
    def c(self):
        return C.byref(self._c)


def factor_f32(pat: SymPattern, val, perturb=False):
    v = np.array(val, dtype=np.float32, copy=True)
    pert = _load().oracle_factor_f32(pat.c, _p(v), int(bool(perturb)))
    return v, float(pert)


def factor_f32_sparse(pat: SymPattern, val, perturb=False):
    """factor_f32 without the dense scatter's signed-zero updates: the same
    bits unless an entry holds -0.0 or a factor is non-finite, at the cost
    of the scalar update count (used for the largest orders)."""
    v = np.array(val, dtype=np.float32, copy=True)
    pert = _load().oracle_factor_f32_sparse(pat.c, _p(v), int(bool(perturb)))
    return v, float(pert)


def conflicts(pat: SymPattern):
    mark = np.zeros(pat.nnz, dtype=np.uint8)
    cnt = _load().oracle_conflicts(pat.c, _p(mark))
    return int(cnt), mark.astype(bool)


def solve_f32(pat: SymPattern, lu, col_perm, row_perm_inv, col_scale_perm, row_scale, rhs):
    lu = np.ascontiguousarray(lu, dtype=np.float32)
    rhs = np.ascontiguousarray(rhs, dtype=np.float32)
    x = np.empty(pat.n, dtype=np.float32)
    args = [np.ascontiguousarray(col_perm, dtype=np.uint32), np.ascontiguousarray(row_perm_inv, dtype=np.uint32),
            np.ascontiguousarray(col_scale_perm, dtype=np.float64), np.ascontiguousarray(row_scale, dtype=np.float64)]
    _load().oracle_solve_f32(pat.c, _p(lu), *[_p(a) for a in args], _p(rhs), _p(x))
    return x


def factor_f64(pat: SymPattern, val):
    v = np.array(val, dtype=np.float64, copy=True)
    _load().oracle_factor_f64(pat.c, _p(v))
    return v


def solve_f64(pat: SymPattern, lu, an, b, transpose=False):
    """an: object with row_perm, row_perm_inv, col_perm, col_perm_inv, col_scale_perm, row_scale."""
    lu = np.ascontiguousarray(lu, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.empty(pat.n, dtype=np.float64)
    arrs = [np.ascontiguousarray(getattr(an, k), dtype=np.uint32)
            for k in ("row_perm", "row_perm_inv", "col_perm", "col_perm_inv")]
    arrs += [np.ascontiguousarray(an.col_scale_perm, dtype=np.float64),
             np.ascontiguousarray(an.row_scale, dtype=np.float64)]
    _load().oracle_solve_f64(pat.c, _p(lu), *[_p(a) for a in arrs], _p(b), _p(x), int(bool(transpose)))
    return x


def lu_values_f64(pat: SymPattern, dump_ax, dump_ai, dump_ap):
    """FP64 placement of the dumped matrix into the LU pattern (predictLU in
    double precision)."""
    n = pat.n
    v = np.zeros(pat.nnz, dtype=np.float64)
    for c in range(n):
        rows = pat.sym_r_idx[pat.sym_c_ptr[c]:pat.sym_c_ptr[c + 1]]
        s, e = int(dump_ap[c]), int(dump_ap[c + 1])
        pos = np.searchsorted(rows, dump_ai[s:e])
        v[pat.sym_c_ptr[c] + pos] = dump_ax[s:e]
    return v


def dense_doolittle(M: np.ndarray):
    """Textbook dense LU without pivoting (kij order) — an independent check
    of the sparse eliminations on small matrices.  Returns L\\U packed."""
    A = np.array(M, dtype=np.float64, copy=True)
    n = A.shape[0]
    for k in range(n):
        A[k + 1:, k] /= A[k, k]
        A[k + 1:, k + 1:] -= np.outer(A[k + 1:, k], A[k, k + 1:])
    return A


def pattern_to_dense(pat: SymPattern, val):
    n = pat.n
    D = np.zeros((n, n), dtype=np.float64)
    for c in range(n):
        s, e = pat.sym_c_ptr[c], pat.sym_c_ptr[c + 1]
        D[pat.sym_r_idx[s:e], c] = val[s:e]
    return D


def refresh(pat: SymPattern, an, indices, indptr, ax0, order=0):
    """Refactor values of ax0 (caller CSC order) with the frozen analysis
    `an` (row_perm, col_perm_inv, row_scale, col_scale_perm): NICSLU's
    refactor scaling, my_DumpA, predictLU.  order 0 = (a * rs) * csp (NICSLU's
    refactor product), 1 = a * (rs * csp) (the analysis product).  Returns
    (dumped FP64 values, FP32 LU-pattern values)."""
    ap = np.ascontiguousarray(indptr, dtype=np.uint32)
    ai = np.ascontiguousarray(indices, dtype=np.uint32)
    ax0 = np.ascontiguousarray(ax0, dtype=np.float64)
    arrs = [np.ascontiguousarray(an.row_perm, dtype=np.uint32), np.ascontiguousarray(an.col_perm_inv, dtype=np.uint32),
            np.ascontiguousarray(an.row_scale, dtype=np.float64),
            np.ascontiguousarray(an.col_scale_perm, dtype=np.float64)]
    dump = np.empty(ax0.size)
    val = np.empty(pat.nnz, dtype=np.float32)
    _load().oracle_refresh(pat.c, ax0.size, _p(ap), _p(ai), _p(ax0), *[_p(a) for a in arrs], int(order),
                           _p(dump), _p(val))
    return dump, val


def abft_cca(pat: SymPattern, val):
    """src/symbolic.cc:163-170: FP32 column sums of the values before the
    factorization."""
    v = np.ascontiguousarray(val, dtype=np.float32)
    cca = np.empty(pat.n, dtype=np.float32)
    _load().oracle_abft_cca(pat.c, _p(v), _p(cca))
    return cca


def abft_check(pat: SymPattern, lu, cca, tol=1e-5):
    """src/symbolic.cc:172-197: (first failing column or -1, the rebuilt
    checksums CCA_ABFT, the largest |CCA_ABFT - CCA|)."""
    lu = np.ascontiguousarray(lu, dtype=np.float32)
    cca = np.ascontiguousarray(cca, dtype=np.float32)
    abft = np.empty(pat.n, dtype=np.float32)
    mx = np.zeros(1, dtype=np.float32)
    first = _load().oracle_abft_check(pat.c, _p(lu), _p(cca), float(tol), _p(abft), _p(mx))
    return int(first), abft, float(mx[0])
