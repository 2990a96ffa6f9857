"""The public drop-in surface on the GPU: pyglu.splu / spsolve /
GLUFactorization (the reference's pyglu/__init__.py:66-122 and
src/glu_binding.cpp:44-155).  FP64 in, FP32 factor and solve, FP64 out: the
result must be the oracle's FP32 factorization and the reference's FP32
solve (restated in the oracle) of the float-cast right-hand side, cast to
FP64, bit for bit."""

import numpy as np
import pytest
import scipy.sparse as sp

from helpers import mna_case

pytestmark = pytest.mark.gpu


def _expected(data, indices, indptr, n, b, perturb):
    from oracle import oracle as O
    from pyglu.analysis import Analysis
    an = Analysis(data, indices, indptr, n)
    pat = O.SymPattern(an)
    lu, _ = O.factor_f32(pat, an.val, perturb=perturb)
    x = O.solve_f32(pat, lu, an.col_perm, an.row_perm_inv, an.col_scale_perm, an.row_scale,
                    b.astype(np.float32))
    return x.astype(np.float64)


@pytest.mark.parametrize("name", ["add32", "small"])
def test_splu_spsolve_bit_exact(name):
    import pyglu
    from pyglu import mna
    d, i, p, n = mna_case(name)
    A = sp.csc_matrix((d, i, p), shape=(n, n))
    b = mna.rhs(n, 17)
    print(f"[{name}] n={n} nnz={d.size}", flush=True)
    for perturb in (False, True):
        want = _expected(d, i, p, n, b, perturb)
        F = pyglu.splu(A, perturb=perturb)
        assert F.n == n
        x = F.solve(b)
        assert x.dtype == np.float64 and x.shape == (n,)
        assert np.array_equal(x.view(np.uint64), want.view(np.uint64)), f"splu(perturb={perturb}).solve"
        x2 = pyglu.spsolve(A, b, perturb=perturb)
        assert np.array_equal(x2.view(np.uint64), want.view(np.uint64)), f"spsolve(perturb={perturb})"
        G = pyglu.GLUFactorization(d, i, p, n, perturb)
        assert np.array_equal(G.solve(b).view(np.uint64), want.view(np.uint64))
        # the 4-tuple input form and an unsorted / duplicated COO input (tocsc,
        # sum_duplicates, sort_indices) give the same answer
        x3 = pyglu.spsolve((d, i, p, (n, n)), b, perturb=perturb)
        assert np.array_equal(x3.view(np.uint64), want.view(np.uint64))
    coo = A.tocoo()
    half = coo.data / 2
    dup = sp.coo_matrix((np.concatenate([half, coo.data - half]),
                         (np.concatenate([coo.row, coo.row]), np.concatenate([coo.col, coo.col]))), shape=(n, n))
    x4 = pyglu.spsolve(dup.tocsr(), b)
    B = dup.tocsc()
    B.sum_duplicates()
    B.sort_indices()
    want4 = _expected(B.data, B.indices, B.indptr, n, b, False)
    assert np.array_equal(x4.view(np.uint64), want4.view(np.uint64))


def test_dropin_errors():
    """Error mapping of src/glu_binding.cpp:51-77 and :131-135: ValueError for
    shapes and a wrong right-hand-side length, TypeError for a dense input,
    RuntimeError when preprocessing fails (structurally singular)."""
    import pyglu
    from pyglu import mna
    d, i, p, n = mna_case("mna600")
    F = pyglu.splu(sp.csc_matrix((d, i, p), shape=(n, n)))
    with pytest.raises(ValueError):
        F.solve(np.ones(n + 1))
    with pytest.raises(ValueError):
        pyglu.spsolve(sp.csc_matrix(np.ones((3, 4))), np.ones(3))
    with pytest.raises(TypeError):
        pyglu.spsolve(np.eye(3), np.ones(3))
    with pytest.raises(RuntimeError):
        pyglu.splu(sp.csc_matrix((np.ones(2), (np.array([0, 1]), np.array([0, 0]))), shape=(2, 2)))
    # a solve after the errors still works on the earlier factorization
    b = mna.rhs(n, 3)
    assert np.array_equal(F.solve(b).view(np.uint64), _expected(d, i, p, n, b, False).view(np.uint64))
