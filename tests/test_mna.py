"""The seeded MNA generator: shapes, structure and reproducibility."""

import numpy as np
import scipy.sparse as sp

from pyglu import mna


def test_shapes_near_targets():
    for spec, seed, n, nnz in [(mna.SPEC_SMALL, 1, 4960, 24_000), (mna.SPEC_BATCH, 3, 20_000, 120_000),
                               (mna.SPEC_GRAD, 5, 50_000, 300_000)]:
        d, i, p, n_ = mna.mna_csc(spec, seed)
        assert n_ == n and abs(d.size - nnz) / nnz < 0.1


def test_structure():
    spec = mna.SPEC_SMALL
    d, i, p, n = mna.mna_csc(spec, 1)
    A = sp.csc_matrix((d, i, p), shape=(n, n))
    S = (A != 0).astype(int)
    assert (S != S.T).nnz == 0                       # structurally symmetric
    diag = A.diagonal()
    nn = spec.n_nodes
    assert np.all(diag[:nn] > 0) and np.all(diag[nn:] == 0)   # vsrc rows: no diagonal
    G = A[:nn, :nn].toarray()
    off = np.abs(G).sum(1) - np.abs(np.diag(G))
    assert np.all(np.diag(G) > off)                  # strictly diagonally dominant node block
    deg = np.diff(A.indptr)[:nn] - 1
    assert deg.max() >= 300                          # rails present
    assert np.all(np.diff(p) > 0) and np.all(np.diff(i[p[0]:p[1]]) > 0)


def test_reproducible_and_seed_sensitive():
    a = mna.mna_csc(mna.SPEC_SMALL, 1)
    b = mna.mna_csc(mna.SPEC_SMALL, 1)
    c = mna.mna_csc(mna.SPEC_SMALL, 2)
    assert all(np.array_equal(x, y) for x, y in zip(a[:3], b[:3]))
    assert not np.array_equal(a[0], c[0]) or not np.array_equal(a[1], c[1])


def test_uneven_specs():
    specs = mna.uneven_specs()
    ns = [s.n for s, _ in specs]
    assert len(specs) == 64 and min(ns) >= 1000 and max(ns) <= 50_000


def test_value_drift():
    d = np.linspace(1, 2, 100)
    v = mna.perturb_values(d, 3000)
    assert np.all(np.abs(v / d - 1) <= 0.1)
