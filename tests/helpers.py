"""Shared helpers for the tests (matrix cases and error metrics)."""

from __future__ import annotations

import os

import numpy as np
import scipy.sparse as sp

from pyglu import mna

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def add32():
    """src/add32.mtx of the reference (n=4960, 23,884 stored entries incl.
    4,036 explicit zeros), as CSC with the explicit zeros kept."""
    z = np.load(os.path.join(GOLDEN, "add32.npz"))
    n = int(z["n"])
    A = sp.csc_matrix((z["val"], (z["row"], z["col"])), shape=(n, n))
    A.sort_indices()
    # coo->csc keeps explicit zeros as stored entries
    return A.data.copy(), A.indices.astype(np.int32), A.indptr.astype(np.int32), n


def mna_case(name):
    specs = {
        "mna600": (mna.MNASpec(n_nodes=600, n_vsrc=8, n_rails=2, rail_deg=40), 21),
        "mna2k": (mna.MNASpec(n_nodes=1980, n_vsrc=20, n_rails=3, rail_deg=120), 7),
        "small": (mna.SPEC_SMALL, mna.SEED_SMALL),
        "batch": (mna.SPEC_BATCH, mna.SEED_BATCH),
        "grad": (mna.SPEC_GRAD, mna.SEED_GRAD),
    }
    if name == "add32":
        return add32()
    spec, seed = specs[name]
    return mna.mna_csc(spec, seed)


def entry_rel_err(got, want, col_ptr):
    """Per-entry relative error |got - want| / max(|want|, 1e-3 * |col|_inf):
    relative to the entry itself, floored at a thousandth of its column's
    largest magnitude (entries that cancelled)."""
    got = np.asarray(got, dtype=np.float64)
    want = np.asarray(want, dtype=np.float64)
    colmax = np.maximum.reduceat(np.abs(want), col_ptr[:-1].astype(np.int64))
    counts = np.diff(col_ptr.astype(np.int64))
    floor = np.repeat(colmax, counts) * 1e-3
    scale = np.maximum(np.abs(want), floor)
    scale[scale == 0] = 1.0
    return np.abs(got - want) / scale


def backward_error(data, indices, indptr, n, x, b):
    A = sp.csc_matrix((data, indices, indptr), shape=(n, n))
    r = b - A @ x
    anorm = np.abs(A).sum(axis=1).max()
    return float(np.abs(r).max() / (anorm * np.abs(x).max() + np.abs(b).max()))


class DensePattern:
    """Symbolic_Matrix arrays for a small dense boolean pattern, computed by
    brute force: dense boolean right-looking elimination for the fill and an
    iterate-to-fixpoint longest path for the levels (the dependency rule of
    src/symbolic.cc:115-130).  Independent of the product's symbolic code."""

    def __init__(self, mask):
        F = np.array(mask, dtype=bool)
        n = F.shape[0]
        np.fill_diagonal(F, True)
        for k in range(n):
            rows = np.nonzero(F[k + 1:, k])[0] + k + 1
            cols = np.nonzero(F[k, k + 1:])[0] + k + 1
            F[np.ix_(rows, cols)] = True
        self.F = F
        self.n = n
        self.sym_c_ptr = np.zeros(n + 1, np.uint32)
        idx, lcp = [], []
        for c in range(n):
            r = np.nonzero(F[:, c])[0]
            lcp.append(self.sym_c_ptr[c] + np.searchsorted(r, c))
            idx.extend(r)
            self.sym_c_ptr[c + 1] = self.sym_c_ptr[c] + r.size
        self.sym_r_idx = np.array(idx, np.uint32)
        self.l_col_ptr = np.array(lcp, np.uint32)
        self.csr_r_ptr = np.zeros(n + 1, np.uint32)
        cidx, cdp = [], []
        for r in range(n):
            c = np.nonzero(F[r, :])[0]
            cdp.append(self.csr_r_ptr[r] + np.searchsorted(c, r))
            cidx.extend(c)
            self.csr_r_ptr[r + 1] = self.csr_r_ptr[r] + c.size
        self.csr_c_idx = np.array(cidx, np.uint32)
        self.csr_diag_ptr = np.array(cdp, np.uint32)
        has_l = np.array([F[c + 1:, c].any() for c in range(n)])
        deps = [[r for r in range(i) if (F[r, i] and has_l[r]) or F[i, r]] for i in range(n)]
        lev = np.zeros(n, int)
        changed = True
        while changed:
            changed = False
            for i in range(n):
                want = 1 + max((lev[d] for d in deps[i]), default=-1)
                if want != lev[i]:
                    lev[i], changed = want, True
        nl = lev.max() + 1
        self.level_ptr = np.concatenate([[0], np.cumsum(np.bincount(lev, minlength=nl))]).astype(np.int32)
        self.level_idx = np.concatenate([np.nonzero(lev == l)[0] for l in range(nl)]).astype(np.int32)
        self.num_lev = nl

    def values(self, dense):
        """Pattern-ordered values of a dense matrix."""
        return np.concatenate([dense[np.nonzero(self.F[:, c])[0], c] for c in range(self.n)])
