"""The C-ABI library: it loads, exports every entry point include/glu_hip.h
declares, and validates arguments and patterns on the host (no GPU
needed for these)."""

import ctypes as C
import os
import re

import numpy as np
import pytest

from helpers import DensePattern, mna_case

from pyglu import _capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "glu_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(glu_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _capi.lib()
    names = declared_functions()
    assert len(names) >= 19
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing
    assert set(names) == set(_capi.EXPORTED)


def test_header_is_plain_c():
    """The boundary header compiles as C99 with no HIP/torch headers."""
    import subprocess
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "t.c")
        with open(src, "w") as f:
            f.write('#include "glu_hip.h"\nint main(void){glu_plan *p = 0; (void)p; return 0;}\n')
        subprocess.check_call(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                               "-c", src, "-o", os.path.join(d, "t.o")])


def test_analysis_get_checks_sizes():
    from pyglu.analysis import Analysis
    d, i, p, n = mna_case("mna600")
    a = Analysis(d, i, p, n)
    buf = np.empty(n + 2, dtype=np.uint32)  # one too many
    rc = _capi.lib().glu_analysis_get(a.handle, _capi.F_SYM_C_PTR, _capi.ptr(buf), buf.nbytes)
    assert rc == -2 and b"size mismatch" in _capi.lib().glu_last_error()


def _pattern_struct(P):
    keep = [np.ascontiguousarray(getattr(P, k), dtype=np.uint32) for k in
            ("sym_c_ptr", "sym_r_idx", "l_col_ptr", "csr_r_ptr", "csr_c_idx", "csr_diag_ptr")]
    keep += [np.ascontiguousarray(P.level_ptr, dtype=np.int32), np.ascontiguousarray(P.level_idx, dtype=np.int32)]
    return keep, _capi.Pattern(keep[2].size, keep[1].size, keep[6].size - 1, *(a.ctypes.data for a in keep))


def _create(P):
    keep, pat = _pattern_struct(P)
    h = C.c_void_p()
    rc = _capi.lib().glu_plan_create(C.byref(pat), None, C.byref(h))
    if rc == 0:
        _capi.lib().glu_plan_destroy(h)
    return rc


def test_plan_rejects_bad_patterns():
    """Patterns that would make the kernels index out of bounds or race are
    refused with GLU_ERR_INVALID before any device work."""
    rng = np.random.default_rng(3)
    M = (rng.random((30, 30)) < 0.1) | np.eye(30, dtype=bool)
    good = DensePattern(M)

    class Bad:
        pass

    def variant(**kw):
        b = Bad()
        for k in ("sym_c_ptr", "sym_r_idx", "l_col_ptr", "csr_r_ptr", "csr_c_idx", "csr_diag_ptr",
                  "level_ptr", "level_idx"):
            setattr(b, k, np.array(getattr(good, k)).copy())
        for k, f in kw.items():
            f(getattr(b, k))
        return b

    def bump(i, v):
        def f(a):
            a[i] = a[i] + 1 if v is None else v
        return f

    cases = [
        variant(sym_r_idx=bump(3, 999)),                    # row out of range
        variant(l_col_ptr=bump(int(np.argmax(np.diff(good.sym_c_ptr.astype(np.int64)) > 1)),
                               None)),                    # diagonal pointer off the diagonal
        variant(csr_c_idx=bump(2, 1)),                     # CSR not the transpose
        variant(level_idx=bump(0, int(good.level_idx[1]))),  # levels not a permutation
    ]
    # a level schedule that breaks a dependency (everything in one level)
    b = variant()
    b.level_ptr = np.array([0, 30], dtype=np.int32)
    b.level_idx = np.arange(30, dtype=np.int32)
    cases.append(b)
    # a pattern not closed under fill: drop one fill entry from both views
    fill = np.argwhere(good.F & ~(M | np.eye(30, dtype=bool)))
    if len(fill):
        i0, j0 = fill[0]
        M2 = good.F.copy()
        M2[i0, j0] = False
        raw = DensePattern.__new__(DensePattern)
        raw.__dict__.update(good.__dict__)
        # rebuild arrays from M2 without re-running the fill
        F = M2
        n = 30
        raw.sym_c_ptr = np.concatenate([[0], np.cumsum(F.sum(0))]).astype(np.uint32)
        raw.sym_r_idx = np.concatenate([np.nonzero(F[:, c])[0] for c in range(n)]).astype(np.uint32)
        raw.l_col_ptr = np.array([raw.sym_c_ptr[c] + np.searchsorted(np.nonzero(F[:, c])[0], c) for c in range(n)],
                                 np.uint32)
        raw.csr_r_ptr = np.concatenate([[0], np.cumsum(F.sum(1))]).astype(np.uint32)
        raw.csr_c_idx = np.concatenate([np.nonzero(F[r, :])[0] for r in range(n)]).astype(np.uint32)
        raw.csr_diag_ptr = np.array([raw.csr_r_ptr[r] + np.searchsorted(np.nonzero(F[r, :])[0], r)
                                     for r in range(n)], np.uint32)
        cases.append(raw)
    for k, c in enumerate(cases):
        assert _create(c) == -7, (k, _capi.lib().glu_last_error())


def test_plan_without_device_fails_cleanly():
    """On a host with no HIP device a valid pattern gives GLU_ERR_DEVICE."""
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a device is present")
    except ImportError:
        pass
    rng = np.random.default_rng(4)
    P = DensePattern((rng.random((20, 20)) < 0.1) | np.eye(20, dtype=bool))
    assert _create(P) == -4


def test_null_arguments():
    L = _capi.lib()
    assert L.glu_factor(None, None, 0, None) == -2
    assert L.glu_solve(None, None, None, None, 0, None) == -2
    assert L.glu_batch_solve(None, 1, None, None, None, None, 1e-10, 5, None, None, 0, None) == -2
    assert L.glu_batch_factor(None, 1, None, None, 0, None) == -2
    assert L.glu_abft_sums(None, 1, None, None, None) == -2
    assert L.glu_abft_check(None, 1, None, None, 1e-5, None, None, None) == -2
    h = C.c_void_p()
    assert L.glu_analyze(0, 0, None, None, None, C.byref(h)) == -9
