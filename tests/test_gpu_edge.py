"""Edge cases of the device path beyond the MNA family: the smallest
matrices, random unsymmetric patterns (dense rows and columns, integer ties,
long rows), and exactly singular matrices, on which the factors with and
without the GESP perturbation (src/numeric.cu:139-140, pivot value from
:490-506) are compared with the oracle (bit for bit where the factors stay
finite)."""

import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_mod():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def _csc(A):
    A = sp.csc_matrix(A)
    A.sum_duplicates()
    A.sort_indices()
    return A.data.astype(np.float64).copy(), A.indices.astype(np.int32), A.indptr.astype(np.int32), A.shape[0]


def _same_bits(got, want):
    """Bit-identical, except that a NaN matches any NaN (the payload of a
    NaN is not part of the arithmetic)."""
    got, want = np.asarray(got), np.asarray(want)
    gn, wn = np.isnan(got), np.isnan(want)
    if not np.array_equal(gn, wn):
        return False
    return np.array_equal(got[~gn].view(np.uint32), want[~wn].view(np.uint32))


def _same_factors(got, want):
    """The factors' agreement with the oracle.  All finite: bit-identical.
    Otherwise the reference's dense scatter (src/numeric.cu:51-89) also adds
    -(0 * U(k,j)) to the entries of sub-column j below row k outside L(:,k),
    which turns them to NaN once U(k,j) is infinite or NaN, where the device
    schedule (which touches only the entries the elimination reaches) keeps
    them: every entry the oracle has finite is then equal on the device (a
    zero may differ in sign), and both sets of factors are non-finite."""
    got, want = np.asarray(got), np.asarray(want)
    fin = np.isfinite(want)
    if fin.all():
        return np.array_equal(got.view(np.uint32), want.view(np.uint32))
    return np.array_equal(got[fin], want[fin]) and not np.isfinite(got).all()


def _random_case(t):
    """Case t of the random family: an unsymmetric pattern of 5-400 rows, a
    dense row and column every fifth case, integer-valued entries every
    third, and a permuted diagonal so that a perfect matching exists."""
    rng = np.random.default_rng(1000 + t)
    n = int(rng.integers(5, 400))
    A = sp.random(n, n, density=float(rng.uniform(0.005, 0.06)), random_state=int(rng.integers(1 << 30)),
                  format="lil")
    if t % 5 == 2:
        k = int(rng.integers(0, n))
        A[k, :] = rng.uniform(-1, 1, n)
        A[:, (k + 1) % n] = rng.uniform(-1, 1, (n, 1))
    A = sp.csc_matrix(A)
    if t % 3 == 0:
        A.data[:] = np.round(A.data * 4) + 1.0
    A = A + sp.csc_matrix((rng.uniform(0.5, 2, n), (rng.permutation(n), np.arange(n))), shape=(n, n))
    return _csc(A)


def _check_factor_and_solve(torch, data, indices, indptr, n, perturb):
    from oracle import oracle as O
    from pyglu.analysis import Analysis
    from pyglu.device import Plan
    an = Analysis(data, indices, indptr, n)
    pat = O.SymPattern(an)
    plan = Plan(an)
    lu = torch.from_numpy(an.val.copy()).cuda()
    plan.factor(lu, perturb=perturb)
    torch.cuda.synchronize()
    got = lu.cpu().numpy()
    want, pert = O.factor_f32(pat, an.val, perturb=perturb)
    assert _same_factors(got, want), (n, perturb)
    if np.isfinite(want).all():
        b = np.random.default_rng(n).uniform(-1, 1, n).astype(np.float32)
        x = plan.solve(lu, torch.from_numpy(b).cuda())
        torch.cuda.synchronize()
        xs = O.solve_f32(pat, want, an.col_perm, an.row_perm_inv, an.col_scale_perm, an.row_scale, b)
        assert _same_bits(x.cpu().numpy(), xs), n
    return an, plan, want, pert


@pytest.mark.parametrize("dense", [[[2.0]], [[0.0, 1.0], [1.0, 0.0]], [[1.0, 2.0], [3.0, 4.0]],
                                   [[4.0, -1.0, 0.0], [-1.0, 4.0, -1.0], [0.0, -1.0, 4.0]]])
def test_tiny_matrices(torch_mod, dense):
    """n = 1, 2, 3 (the 2x2 anti-diagonal needs the MC64 row permutation):
    factors and solve bit-identical to the oracle, refinement to 1e-10, and
    a batch of two value sets equal to the single path."""
    torch = torch_mod
    data, indices, indptr, n = _csc(np.array(dense))
    an, plan, _, _ = _check_factor_and_solve(torch, data, indices, indptr, n, False)
    ax = torch.from_numpy(data).cuda()
    lu = plan.refresh(ax).reshape(-1)
    plan.factor(lu)
    b = torch.from_numpy(np.arange(1.0, n + 1.0)).cuda()
    x, it, eta = plan.refine(ax, lu, b, tol=1e-10, maxit=10)
    want = np.linalg.solve(np.array(dense), np.arange(1.0, n + 1.0))
    assert eta <= 1e-10 and np.allclose(x.cpu().numpy(), want, rtol=1e-12, atol=1e-14)
    axd = torch.stack([ax, ax * 1.5]).contiguous()
    bd = torch.stack([b, b]).contiguous()
    lub = plan.batch_factor(axd)
    torch.cuda.synchronize()
    assert torch.equal(lub[0].view(torch.int32), lu.view(torch.int32))
    xb, _, itb, etab = plan.batch_solve(axd, bd, tol=1e-10, maxit=10)
    torch.cuda.synchronize()
    assert float(etab.max()) <= 1e-10
    assert np.allclose(xb[1].cpu().numpy(), want / 1.5, rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("t", range(12))
@pytest.mark.parametrize("perturb", [False, True])
def test_random_patterns(torch_mod, t, perturb):
    """Random unsymmetric patterns: factors (with and without GESP) and the
    exact solve bit-identical to the oracle; the batched refactor of three
    drifted value sets equal to the single path."""
    torch = torch_mod
    data, indices, indptr, n = _random_case(t)
    print(f"[random {t}] n={n} nnz={data.size} perturb={perturb}", flush=True)
    an, plan, _, _ = _check_factor_and_solve(torch, data, indices, indptr, n, perturb)
    rng = np.random.default_rng(t)
    axd = torch.from_numpy(np.stack([data * (1 + rng.uniform(-0.1, 0.1, data.size)) for _ in range(3)])).cuda()
    lub = plan.batch_factor(axd, perturb=perturb)
    for b in range(3):
        one = plan.refresh(axd[b]).reshape(-1)
        plan.factor(one, perturb=perturb)
        torch.cuda.synchronize()
        assert _same_bits(one.cpu().numpy(), lub[b].cpu().numpy()), b


def test_singular_matrix_gesp(torch_mod):
    """An exactly singular matrix (two equal columns): without the
    perturbation a pivot cancels to zero and the factors carry the same
    infinities and NaNs as the oracle's, and the refinement reports a
    non-finite backward error (an error status for NaN); with it the zero
    pivot is replaced by the GESP value, bit for bit as in the oracle."""
    from pyglu._capi import GluError
    torch = torch_mod
    rng = np.random.default_rng(5)
    n = 40
    A = sp.lil_matrix(sp.random(n, n, density=0.08, random_state=3, format="lil") +
                      sp.diags(rng.uniform(0.5, 2, n)))
    A[:, 7] = A[:, 3]
    data, indices, indptr, n = _csc(A)
    an, plan, want, _ = _check_factor_and_solve(torch, data, indices, indptr, n, False)
    assert not np.isfinite(want).all()
    ax = torch.from_numpy(data).cuda()
    lu = plan.refresh(ax).reshape(-1)
    plan.factor(lu)
    try:
        _, _, eta = plan.refine(ax, lu, torch.ones(n, dtype=torch.float64, device="cuda"), tol=1e-10, maxit=3)
        assert not np.isfinite(eta)
    except GluError:
        pass
    _, _, want_p, pert = _check_factor_and_solve(torch, data, indices, indptr, n, True)
    from oracle import oracle as O
    pat = O.SymPattern(an)
    diag = want_p[pat.l_col_ptr]
    assert np.isfinite(want_p).all() and pert > 0 and np.count_nonzero(diag == np.float32(pert)) >= 1
