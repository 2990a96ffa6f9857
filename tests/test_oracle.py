"""Pinning the oracle (CPU): each restated function is checked against
something the mathematics or the reference itself fixes, chosen so that a
plausible mistake (dropped term, wrong sign, transposed operand, wrong
update order, missing GESP quirk) fails at least one check."""

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as sla

from helpers import DensePattern, mna_case

from oracle import oracle as O
from pyglu.analysis import Analysis


def f32(x):
    return np.float32(x)


def test_factor_diagonal_is_identity_op():
    """Diagonal A: no update, no division — factors equal A bit for bit."""
    P = DensePattern(np.eye(6))
    v = np.array([3, -2, 0.5, 7, 1e-3, 9], np.float32)
    lu, _ = O.factor_f32(O.SymPattern(P), v)
    assert np.array_equal(lu, v)


def test_factor_2x2_closed_form():
    """[[a, b], [c, d]]: l = fl(c/a), u22 = fl(d - fl(l*b)) exactly."""
    P = DensePattern(np.ones((2, 2)))
    a, b, c, d = f32(1.7), f32(-0.3), f32(0.9), f32(2.1)
    vals = P.values(np.array([[a, b], [c, d]], np.float32)).astype(np.float32)
    lu, _ = O.factor_f32(O.SymPattern(P), vals)
    l = f32(c / a)
    u22 = f32(d - f32(l * b))
    assert lu.tolist() == [a, l, b, u22]


def test_factor_tridiagonal_recurrence():
    """Tridiagonal: u_{i+1} = d_{i+1} - fl(fl(s_i/u_i) * t_i), no fill."""
    n = 40
    rng = np.random.default_rng(4)
    d = rng.uniform(2, 3, n).astype(np.float32)
    s = rng.uniform(-1, 1, n - 1).astype(np.float32)   # sub-diagonal
    t = rng.uniform(-1, 1, n - 1).astype(np.float32)   # super-diagonal
    D = np.diag(d) + np.diag(s, -1) + np.diag(t, 1)
    P = DensePattern(D != 0)
    assert P.sym_r_idx.size == 3 * n - 2
    lu, _ = O.factor_f32(O.SymPattern(P), P.values(D).astype(np.float32))
    got = O.pattern_to_dense(O.SymPattern(P), lu)
    u = d[0]
    assert got[0, 0] == u
    for i in range(n - 1):
        l = f32(s[i] / u)
        assert got[i + 1, i] == l and got[i, i + 1] == t[i]
        u = f32(d[i + 1] - f32(l * t[i]))
        assert got[i + 1, i + 1] == u


def test_factor_upper_triangular_unchanged():
    """Upper-triangular A: empty L everywhere, the factors are A itself."""
    rng = np.random.default_rng(5)
    D = np.triu(rng.uniform(-1, 1, (12, 12)) * (rng.random((12, 12)) < 0.4)) + np.eye(12) * 3
    P = DensePattern(D != 0)
    v = P.values(D).astype(np.float32)
    lu, _ = O.factor_f32(O.SymPattern(P), v)
    assert np.array_equal(lu, v)


@pytest.mark.parametrize("seed", range(6))
def test_factor_f64_equals_dense_doolittle(seed):
    """Sparse FP64 elimination = textbook dense Doolittle on random
    diagonally dominant patterns (fill included), to rounding."""
    rng = np.random.default_rng(seed)
    n = int(rng.integers(10, 60))
    M = rng.uniform(-1, 1, (n, n)) * (rng.random((n, n)) < 0.12)
    M += np.diag(np.abs(M).sum(1) + 1)
    P = DensePattern(M != 0)
    sp_ = O.SymPattern(P)
    lu = O.factor_f64(sp_, P.values(M))
    D = O.dense_doolittle(M)
    got = O.pattern_to_dense(sp_, lu)
    assert np.allclose(got, D, rtol=1e-12, atol=1e-13 * np.abs(D).max())
    assert np.all(D[~P.F] == 0)  # the brute-force fill pattern is closed


def test_factor_on_analysed_mna_vs_dense():
    """On a real preprocessed MNA matrix (permuted, scaled, with MC64 pivots):
    FP64 oracle = dense Doolittle; FP32 oracle within FP32 accuracy of it."""
    d, i, p, n = mna_case("mna600")
    an = Analysis(d, i, p, n)
    pat = O.SymPattern(an)
    v64 = O.lu_values_f64(pat, an.dump_ax, an.dump_ai, an.dump_ap)
    assert np.array_equal(v64.astype(np.float32), an.val)   # predictLU is the FP32 cast
    lu64 = O.factor_f64(pat, v64)
    D = O.dense_doolittle(O.pattern_to_dense(pat, v64))
    assert np.abs(O.pattern_to_dense(pat, lu64) - D).max() <= 1e-12 * np.abs(D).max()
    lu32, _ = O.factor_f32(pat, an.val)
    # |L||U| bounds the perturbation of each factor entry (componentwise)
    absL = np.tril(np.abs(D), -1) + np.eye(n)
    absU = np.triu(np.abs(D))
    S = O.pattern_to_dense(pat, np.ones(pat.nnz)) * (absL @ absU)
    S_entries = np.concatenate([S[pat.sym_r_idx[pat.sym_c_ptr[c]:pat.sym_c_ptr[c + 1]], c] for c in range(n)])
    diag = np.repeat(np.abs(lu64[pat.l_col_ptr]), np.diff(pat.sym_c_ptr.astype(np.int64)))
    rows = pat.sym_r_idx.astype(np.int64)
    cols = np.repeat(np.arange(n), np.diff(pat.sym_c_ptr.astype(np.int64)))
    scale = np.where(rows > cols, S_entries / diag, S_entries)
    assert np.max(np.abs(lu32 - lu64) / np.maximum(np.abs(lu64), scale)) < 1e-4


def test_abft_identity():
    """src/symbolic.cc:163-197: column sums of A equal the checksum rebuilt
    from the factors, CCA_j = sum_i (1 + sum_{r>i} L_ri) U_ij."""
    d, i, p, n = mna_case("mna600")
    an = Analysis(d, i, p, n)
    pat = O.SymPattern(an)
    v64 = O.lu_values_f64(pat, an.dump_ax, an.dump_ai, an.dump_ap)
    lu = O.factor_f64(pat, v64)
    cca = np.add.reduceat(v64, pat.sym_c_ptr[:-1].astype(np.int64))
    ccl = np.ones(n)
    for c in range(n):
        ccl[c] += lu[pat.l_col_ptr[c] + 1:pat.sym_c_ptr[c + 1]].sum()
    abft = np.zeros(n)
    for c in range(n):
        s, e = pat.sym_c_ptr[c], pat.l_col_ptr[c] + 1
        abft[c] = (ccl[pat.sym_r_idx[s:e]] * lu[s:e]).sum()
    assert np.allclose(abft, cca, rtol=1e-10, atol=1e-10 * np.abs(cca).max())


def test_planted_mistakes_are_caught():
    """A dropped update, a sign flip or an L/U transposition in the factors
    is caught by the dense Doolittle check used above."""
    rng = np.random.default_rng(11)
    n = 30
    M = rng.uniform(-1, 1, (n, n)) * (rng.random((n, n)) < 0.15)
    M += np.diag(np.abs(M).sum(1) + 1)
    P = DensePattern(M != 0)
    sp_ = O.SymPattern(P)
    D = O.dense_doolittle(M)
    good = O.factor_f64(sp_, P.values(M))

    def ok(lu):
        return np.allclose(O.pattern_to_dense(sp_, lu), D, rtol=1e-12, atol=1e-13)

    assert ok(good)
    off = np.nonzero(P.sym_r_idx != np.repeat(np.arange(n), np.diff(P.sym_c_ptr.astype(np.int64))))[0]
    bad = good.copy()
    bad[off[len(off) // 2]] *= -1
    assert not ok(bad)
    # a "dropped term": the elimination without column k's contributions
    k = int(np.argmax(np.diff(P.csr_r_ptr.astype(np.int64)) > 1))
    Mk = M.copy()
    Mk[k + 1:, k] = 0
    assert not ok(O.factor_f64(sp_, P.values(Mk)))
    # a transposed operand: eliminating M' instead of M
    assert not ok(O.factor_f64(sp_, P.values(M.T * (P.F != 0))))


def test_gesp_quirks():
    """numeric.cu:133-146 + :490-506: pert = 3.45e-4 * max column |sum|
    (FP32); a tiny pivot is replaced by +pert (sign ignored) only when its
    column has a non-empty L part; a column with empty L is left alone."""
    # 3x3: column 0 pivot tiny negative with L below; column 2 (last, empty L) zero pivot
    M = np.array([[-1e-9, 1.0, 0.0],
                  [2.0, 3.0, 1.0],
                  [0.0, 1.0, 0.0]], np.float32)
    P = DensePattern(M != 0)
    # make (2,2) structurally present so the last pivot exists with value 0
    v = P.values(M).astype(np.float32)
    sp_ = O.SymPattern(P)
    lu, pert = O.factor_f32(sp_, v, perturb=True)
    colsum = max(float(np.abs(M[:, c]).sum(dtype=np.float32)) for c in range(3))
    assert pert == f32(f32(3.45e-4) * f32(colsum))
    D = O.pattern_to_dense(sp_, lu)
    assert D[0, 0] == pert                      # -1e-9 -> +pert
    assert D[1, 0] == f32(f32(2.0) / f32(pert))
    # no perturbation: the tiny pivot stays
    lu0, pert0 = O.factor_f32(sp_, v, perturb=False)
    assert pert0 == 0 and O.pattern_to_dense(sp_, lu0)[0, 0] == f32(-1e-9)
    # last column has empty L: never perturbed, even when its pivot is tiny
    M2 = np.array([[1.0, 0.0], [0.0, 1e-12]], np.float32)
    P2 = DensePattern(M2 != 0)
    lu2, pert2 = O.factor_f32(O.SymPattern(P2), P2.values(M2).astype(np.float32), perturb=True)
    assert pert2 > 1e-12 and lu2[-1] == f32(1e-12)


@pytest.mark.parametrize("name", ["mna600", "add32"])
def test_solve_f32_bit_exact_with_reference(name):
    """oracle_solve_f32 == the reference's Symbolic_Matrix::solve compiled in
    oracle/_ref (same factors, same rhs)."""
    from oracle import ref
    if not ref.available():
        pytest.skip("oracle/_ref not built")
    d, i, p, n = mna_case(name)
    r = ref.analyze(d, i, p, n)
    pat = O.SymPattern(r)
    lu, _ = O.factor_f32(pat, r.val)
    rhs = np.random.default_rng(2).uniform(-1, 1, n).astype(np.float32)
    x_ref = r.solve(lu, rhs)
    x = O.solve_f32(pat, lu, r.col_perm, r.row_perm_inv, r.col_scale_perm, r.row_scale, rhs)
    assert np.array_equal(x.view(np.uint32), x_ref.view(np.uint32))


def test_golden_oracle_and_solve_digests():
    """The oracle's FP32 factors and the reference solve on them, as recorded
    by the reference host code (tests/golden/analysis_sha256.json)."""
    import hashlib
    import json
    import os

    from helpers import GOLDEN
    from pyglu import mna
    with open(os.path.join(GOLDEN, "analysis_sha256.json")) as f:
        gold = json.load(f)
    for name, case in [("add32", mna_case("add32")), ("mna_small", mna_case("small"))]:
        d, i, p, n = case
        an = Analysis(d, i, p, n)
        pat = O.SymPattern(an)
        lu, _ = O.factor_f32(pat, an.val)
        assert hashlib.sha256(lu.tobytes()).hexdigest() == gold[name]["oracle_lu32"]
        rhs = mna.rhs(n, 77).astype(np.float32)
        x = O.solve_f32(pat, lu, an.col_perm, an.row_perm_inv, an.col_scale_perm, an.row_scale, rhs)
        assert hashlib.sha256(x.tobytes()).hexdigest() == gold[name]["ref_solve_f32"]


@pytest.mark.parametrize("transpose", [False, True])
def test_solve_f64_matches_scipy(transpose):
    d, i, p, n = mna_case("mna2k")
    an = Analysis(d, i, p, n)
    pat = O.SymPattern(an)
    lu = O.factor_f64(pat, O.lu_values_f64(pat, an.dump_ax, an.dump_ai, an.dump_ap))
    b = np.random.default_rng(8).uniform(-1, 1, n)
    A = sp.csc_matrix((d, i, p), shape=(n, n))
    want = sla.spsolve((A.T if transpose else A).tocsc(), b)
    got = O.solve_f64(pat, lu, an, b, transpose=transpose)
    assert np.abs(got - want).max() <= 1e-10 * np.abs(want).max()


def test_conflicts_brute_force():
    """oracle_conflicts marks exactly the entries two columns of one level
    both update (brute force on the dense pattern)."""
    rng = np.random.default_rng(12)
    M = (rng.random((40, 40)) < 0.08) | np.eye(40, dtype=bool)
    P = DensePattern(M)
    sp_ = O.SymPattern(P)
    _, mark = O.conflicts(sp_)
    F = P.F
    lev = np.zeros(P.n, int)
    for l in range(P.num_lev):
        lev[P.level_idx[P.level_ptr[l]:P.level_ptr[l + 1]]] = l
    cnt = {}
    for k in range(P.n):
        for i in np.nonzero(F[k + 1:, k])[0] + k + 1:
            for j in np.nonzero(F[k, k + 1:])[0] + k + 1:
                cnt[(i, j, lev[k])] = cnt.get((i, j, lev[k]), 0) + 1
    want = np.zeros_like(mark)
    pos = {}
    for c in range(P.n):
        for q in range(P.sym_c_ptr[c], P.sym_c_ptr[c + 1]):
            pos[(P.sym_r_idx[q], c)] = q
    for (i, j, l), c in cnt.items():
        if c >= 2:
            want[pos[(i, j)]] = True
    assert np.array_equal(mark, want)


def test_abft_closed_forms():
    """ABFT on a diagonal matrix (CCL = 1, checksums = the pivots) and on a
    2x2 matrix (CCL_0 = 1 + c/a: (1 + l) a = a + c, (1 + l) b + (d - l b) =
    b + d) — src/symbolic.cc:163-197."""
    M = np.diag(np.array([2.0, -3.0, 0.5, 7.0], np.float32))
    P = DensePattern(M != 0)
    pat = O.SymPattern(P)
    v = P.values(M).astype(np.float32)
    lu, _ = O.factor_f32(pat, v)
    cca = O.abft_cca(pat, v)
    assert np.array_equal(cca, np.array([2.0, -3.0, 0.5, 7.0], np.float32))
    first, abft, mx = O.abft_check(pat, lu, cca)
    assert first == -1 and mx == 0.0 and np.array_equal(abft, cca)
    a, b, c, d = 4.0, 1.0, 2.0, 3.0
    M2 = np.array([[a, b], [c, d]], np.float32)
    P2 = DensePattern(M2 != 0)
    pat2 = O.SymPattern(P2)
    v2 = P2.values(M2).astype(np.float32)
    lu2, _ = O.factor_f32(pat2, v2)
    cca2 = O.abft_cca(pat2, v2)
    first2, abft2, mx2 = O.abft_check(pat2, lu2, cca2)
    assert np.array_equal(cca2, np.array([a + c, b + d], np.float32))
    assert first2 == -1 and np.allclose(abft2, [a + c, b + d], rtol=1e-6)
    # a wrong pivot shows up in its own column
    bad = lu2.copy()
    bad[pat2.l_col_ptr[1]] += 0.5
    assert O.abft_check(pat2, bad, cca2)[0] == 1


@pytest.mark.parametrize("name", ["mna600", "add32"])
def test_abft_matches_reference(name):
    """oracle_abft_* against the reference's own ABFTCalculateCCA /
    ABFTCheckResult (oracle/_ref, GLU_DEBUG=1): CCA bit for bit, the same
    verdict on the oracle's factors and the same first failing column after
    planted errors in two columns."""
    from oracle import ref
    if not ref.available():
        pytest.skip("oracle/_ref not built")
    d, i, p, n = mna_case(name)
    r = ref.analyze(d, i, p, n)
    pat = O.SymPattern(r)
    lu, _ = O.factor_f32(pat, r.val)
    cca_ref, first_ref = r.abft(r.val, lu)
    cca = O.abft_cca(pat, r.val)
    assert np.array_equal(cca.view(np.uint32), cca_ref.view(np.uint32))
    first, abft, mx = O.abft_check(pat, lu, cca)
    assert first == first_ref == -1
    bad = lu.copy()
    c1, c2 = n // 3, (2 * n) // 3
    bad[pat.l_col_ptr[c2]] += 1.0          # pivot of column c2
    bad[pat.sym_c_ptr[c1]] += 1.0          # first (U or pivot) entry of column c1
    first_b, _, _ = O.abft_check(pat, bad, cca)
    assert first_b == r.abft(r.val, bad)[1] == c1


def test_refresh_golden_digests():
    """oracle_refresh on a drifted value set of every golden input equals the
    reference's own refactor values (NICSLU's _I_NicsLU_MC64ScaleForRefact,
    my_DumpA, predictLU), recorded as digests by tests/golden/make_golden.py:
    the dumped FP64 values and the FP32 LU-pattern values, bit for bit."""
    import hashlib

    from helpers import REFACT_SEED, golden_digests, golden_inputs
    from pyglu import mna
    gold = golden_digests()
    sha = lambda a: hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()
    for name, (d, i, p, n) in golden_inputs():
        an = Analysis(d, i, p, n)
        dump, val = O.refresh(O.SymPattern(an), an, i, p, mna.perturb_values(d, REFACT_SEED))
        assert sha(dump) == gold[name]["refact_dump_ax"], name
        assert sha(val) == gold[name]["refact_val"], name


@pytest.mark.parametrize("name", ["mna600", "add32", "mna2k", "small", "unsym"])
def test_refresh_bit_exact_with_reference(name):
    """Live against the reference build: drifted value sets and a value set
    on which the two product orders round to different FP32 values.  On the
    latter, the analysis order a * (rs * csp) (order=1) must differ from the
    reference in FP32, so the check tells the two orders apart; and with
    order=1 the unchanged values reproduce the analysis's own dump and
    predictLU values bit for bit (both are that product)."""
    from helpers import order_sensitive_values, unsym_case
    from oracle import ref
    from pyglu import mna
    if not ref.available():
        pytest.skip("oracle/_ref not built")
    d, i, p, n = unsym_case() if name == "unsym" else mna_case(name)
    r = ref.analyze(d, i, p, n)
    pat = O.SymPattern(r)
    adv, sensitive = order_sensitive_values(r, i, p, d)
    for ax0 in [mna.perturb_values(d, 7100 + s) for s in range(3)] + [adv]:
        dump_ref, val_ref = r.refact_values(ax0)
        dump, val = O.refresh(pat, r, i, p, ax0)
        assert np.array_equal(dump.view(np.uint64), dump_ref.view(np.uint64))
        assert np.array_equal(val.view(np.uint32), val_ref.view(np.uint32))
    if name != "add32":   # add32's column scales are all powers of two: both orders agree
        _, val1 = O.refresh(pat, r, i, p, adv, order=1)
        assert sensitive.sum() >= (100 if name == "unsym" else 1)
        assert np.count_nonzero(val1.view(np.uint32) != val_ref.view(np.uint32)) == sensitive.sum()
    dump1, val1 = O.refresh(pat, r, i, p, d, order=1)
    assert np.array_equal(dump1.view(np.uint64), r.dump_ax.view(np.uint64))
    assert np.array_equal(val1.view(np.uint32), r.val.view(np.uint32))


@pytest.mark.parametrize("name", ["mna600", "mna2k", "small", "add32", "batch"])
def test_factor_sparse_walk_equals_dense_scatter(name):
    """oracle_factor_f32_sparse (updates of L(:,k)'s rows only) equals the
    dense-scatter restatement bit for bit, with and without GESP, on the
    analysed matrices and on random patterns with dense rows and ties."""
    d, i, p, n = mna_case(name)
    an = Analysis(d, i, p, n)
    pat = O.SymPattern(an)
    for perturb in (False, True):
        a, pa = O.factor_f32(pat, an.val, perturb=perturb)
        b, pb = O.factor_f32_sparse(pat, an.val, perturb=perturb)
        assert pa == pb and np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_factor_sparse_walk_random_patterns():
    rng = np.random.default_rng(77)
    for t in range(12):
        n = int(rng.integers(5, 120))
        M = rng.uniform(-1, 1, (n, n)) * (rng.random((n, n)) < rng.uniform(0.03, 0.2))
        if t % 3 == 0:
            M[int(rng.integers(n)), :] = rng.uniform(-1, 1, n)
        M += np.diag(rng.uniform(0.5, 2.0, n) * rng.choice([-1, 1], n))
        P = DensePattern(M != 0)
        sp_ = O.SymPattern(P)
        v = P.values(M).astype(np.float32)
        for perturb in (False, True):
            a, _ = O.factor_f32(sp_, v, perturb=perturb)
            b, _ = O.factor_f32_sparse(sp_, v, perturb=perturb)
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (t, perturb)
