"""Host analysis (MC64 + AMD + permute/scale + symbolic) is bit-exact with
the reference's host code: against golden digests written by the reference
itself (tests/golden/make_golden.py) and, where oracle/_ref is built, live
against NICSLU + preprocess.c + symbolic.cc on many more inputs."""

import hashlib

import numpy as np
import pytest
import scipy.sparse as sp

from helpers import golden_digests, golden_inputs, mna_case

from pyglu import mna
from pyglu._capi import GluError
from pyglu.analysis import Analysis

ARRAYS = ["row_perm", "row_perm_inv", "col_perm", "col_perm_inv", "row_scale", "col_scale_perm",
          "dump_ax", "dump_ai", "dump_ap", "sym_c_ptr", "sym_r_idx", "l_col_ptr", "csr_r_ptr",
          "csr_c_idx", "csr_diag_ptr", "val", "level_ptr", "level_idx"]


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


GOLD = golden_digests()


def check_golden(name, case):
    d, i, p, n = case
    rec = GOLD[name]
    assert rec["inputs_sha256"] == sha(np.concatenate([d.view(np.uint8), i.astype(np.int32).view(np.uint8),
                                                       p.astype(np.int32).view(np.uint8)]))
    a = Analysis(d, i, p, n)
    assert (a.nnz_lu, a.num_lev) == (rec["nnz_lu"], rec["num_lev"])
    bad = [k for k in ARRAYS if sha(getattr(a, k)) != rec[k]]
    assert not bad, f"{name}: arrays differ from the reference: {bad}"


@pytest.mark.parametrize("name,case", list(golden_inputs()), ids=lambda v: v if isinstance(v, str) else "")
def test_analysis_matches_reference_golden(name, case):
    check_golden(name, case)


@pytest.mark.gpu
def test_analysis_golden_on_device_host():
    """The same golden comparison, run where the GPU tests run: MC64's scale
    factors come from the host's libm exp/log, so the digests are checked on
    the machine that drives the GPU as well."""
    for name, case in golden_inputs():
        check_golden(name, case)


def _ref():
    from oracle import ref
    if not ref.available():
        pytest.skip("oracle/_ref not built (reference sources absent)")
    return ref


def _compare_live(ref, d, i, p, n):
    r = ref.analyze(d, i, p, n)
    a = Analysis(d, i, p, n)
    for k in ARRAYS:
        x, y = getattr(r, k), getattr(a, k)
        assert x.dtype == y.dtype and np.array_equal(x.view(np.uint8), y.view(np.uint8)), k


@pytest.mark.parametrize("name", ["small", "batch"])
def test_analysis_live_mna(name):
    ref = _ref()
    _compare_live(ref, *mna_case(name))


def test_analysis_live_random_stress():
    """Random unsymmetric patterns with ties, dense rows/columns (AMD's dense
    path) and integer-valued entries: MC64's Dijkstra phase, AMD's
    supervariables, absorption and garbage collection all get exercised."""
    ref = _ref()
    rng = np.random.default_rng(99)
    for t in range(40):
        n = int(rng.integers(5, 500))
        A = sp.random(n, n, density=float(rng.uniform(0.004, 0.08)), random_state=int(rng.integers(1 << 30)),
                      format="lil")
        if t % 5 == 2 and n > 20:
            k = int(rng.integers(0, n))
            A[k, :] = rng.uniform(-1, 1, n)
            A[:, (k + 1) % n] = rng.uniform(-1, 1, (n, 1))
        A = sp.csc_matrix(A)
        if t % 3 == 0:
            A.data[:] = 1.0
        if t % 7 == 3:
            A.data = np.round(A.data * 4) / 4 + 0.25
        A = A + sp.csc_matrix((rng.uniform(0.5, 2, n), (rng.permutation(n), np.arange(n))), shape=(n, n))
        A = sp.csc_matrix(A)
        A.sum_duplicates()
        A.sort_indices()
        _compare_live(ref, A.data, A.indices, A.indptr, n)


def test_analysis_unsorted_input_live():
    """Unsorted row indices inside a column (the raw GLUFactorization path
    does not sort) give the same result as the reference."""
    ref = _ref()
    d, i, p, n = mna_case("mna600")
    rng = np.random.default_rng(1)
    d, i = d.copy(), i.copy()
    for c in range(n):
        s, e = p[c], p[c + 1]
        perm = rng.permutation(e - s)
        d[s:e], i[s:e] = d[s:e][perm], i[s:e][perm]
    _compare_live(ref, d, i, p, n)


def _code(fn):
    try:
        fn()
    except GluError as e:
        return e.status
    return 0


def test_error_codes_match_reference():
    """Structurally singular, malformed and duplicated inputs fail with
    NICSLU's codes (nicslu.h:11-28), as the reference analysis does."""
    n = 4
    # column 2 empty -> no perfect matching
    sing = (np.array([1.0, 2, 3]), np.array([0, 1, 3], dtype=np.int32), np.array([0, 1, 2, 2, 3], dtype=np.int32))
    assert _code(lambda: Analysis(*sing, n)) == -5
    # duplicated (adjacent) entry
    dup = (np.array([1.0, 2, 3, 4, 5]), np.array([0, 0, 1, 2, 3], dtype=np.int32),
           np.array([0, 2, 3, 4, 5], dtype=np.int32))
    assert _code(lambda: Analysis(*dup, n)) == -8
    # row index out of range
    bad = (np.array([1.0, 2, 3, 4]), np.array([0, 1, 9, 3], dtype=np.int32), np.array([0, 1, 2, 3, 4], dtype=np.int32))
    assert _code(lambda: Analysis(*bad, n)) == -7
    # empty
    assert _code(lambda: Analysis(np.zeros(0), np.zeros(0, np.int32), np.zeros(n + 1, np.int32), n)) == -9
    from oracle import ref
    if ref.available():
        for case, want in ((sing, -5), (dup, -8), (bad, -7)):
            with pytest.raises(ref.RefError) as ei:
                ref.analyze(*case, n)
            assert ei.value.code == want


def test_python_frontend_validation():
    """pyglu argument errors mirror pyglu/__init__.py:30-63 and
    src/glu_binding.cpp:51-77 (ValueError / TypeError / RuntimeError)."""
    import pyglu
    with pytest.raises(ValueError):
        pyglu.splu(sp.csc_matrix(np.ones((3, 4))))
    with pytest.raises(TypeError):
        pyglu.splu(np.eye(3))
    with pytest.raises(ValueError):
        pyglu.splu((np.ones(2), np.array([0, 1]), np.array([0, 1]), (2, 2)))
    with pytest.raises(ValueError):
        pyglu.GLUFactorization(np.ones(1), np.array([0]), np.array([0, 1]), 0)
    with pytest.raises(ValueError):
        pyglu.GLUFactorization(np.ones(2), np.array([0, -1]), np.array([0, 1, 2]), 2)
    with pytest.raises(ValueError):
        pyglu.GLUFactorization(np.ones(2), np.array([0]), np.array([0, 1, 2]), 2)
    with pytest.raises(RuntimeError):  # structurally singular: preprocessing fails
        pyglu.GLUFactorization(np.ones(2), np.array([0, 1], dtype=np.int32), np.array([0, 2, 2], dtype=np.int32), 2)


def test_analyze_many_matches_single():
    """glu_analyze_many on host threads: every array of every analysis equals
    the one-at-a-time glu_analyze result (12 uneven MNA patterns, 4 threads),
    and a failing member is reported with its index."""
    from pyglu.analysis import analyze_many
    mats = [mna.mna_csc(mna.MNASpec(n_nodes=min(spec.n_nodes, 4000), n_vsrc=min(spec.n_vsrc, 20),
                                    extra_hi=spec.extra_hi, n_rails=spec.n_rails,
                                    rail_deg=min(spec.rail_deg, 100)), seed)
            for spec, seed in mna.uneven_specs(12, seed0=4200)]
    many = analyze_many(mats, threads=4)
    for m, a in zip(mats, many):
        one = Analysis(*m)
        assert (a.n, a.nnz_lu, a.num_lev) == (one.n, one.nnz_lu, one.num_lev)
        for key in ("sym_c_ptr", "sym_r_idx", "l_col_ptr", "level_ptr", "level_idx", "row_perm", "col_perm"):
            assert np.array_equal(getattr(a, key), getattr(one, key)), key
        assert np.array_equal(a.val.view(np.uint32), one.val.view(np.uint32))
        assert np.array_equal(a.row_scale, one.row_scale) and np.array_equal(a.col_scale_perm, one.col_scale_perm)
    # a structurally singular member (an empty column) fails with its index
    d, i, p, n = mats[1]
    col = 3
    keep = np.ones(d.size, bool)
    keep[p[col]:p[col + 1]] = False
    p2 = np.concatenate([[0], np.cumsum([(p[c + 1] - p[c]) if c != col else 0 for c in range(n)])])
    bad = (d[keep], i[keep], p2, n)
    with pytest.raises(GluError, match="matrix 1"):
        analyze_many([mats[0], bad, mats[2]], threads=2)
