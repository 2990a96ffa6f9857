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


def golden_inputs():
    """The inputs of tests/golden/make_golden.py, same order: add32, two MNA
    matrices and eight random patterns (half with equal magnitudes, so that
    MC64/AMD tie-breaks decide everything)."""
    yield "add32", add32()
    yield "mna_small", mna.mna_csc(mna.SPEC_SMALL, mna.SEED_SMALL)
    yield "mna_2k", mna.mna_csc(mna.MNASpec(n_nodes=1980, n_vsrc=20, n_rails=3, rail_deg=120), 7)
    rng = np.random.default_rng(2024)
    for t in range(8):
        n = int(rng.integers(20, 300))
        A = sp.random(n, n, density=float(rng.uniform(0.01, 0.06)), random_state=int(rng.integers(1 << 30)),
                      format="csc")
        if t % 2 == 0:
            A.data[:] = 1.0
        A = A + sp.csc_matrix((np.ones(n), (rng.permutation(n), np.arange(n))), shape=(n, n))
        A = sp.csc_matrix(A)
        A.sum_duplicates()
        A.sort_indices()
        yield f"random{t}", (A.data.copy(), A.indices.astype(np.int32), A.indptr.astype(np.int32), n)


def golden_digests():
    """tests/golden/analysis_sha256.json: per input, SHA-256 digests of the
    reference's analysis arrays, solve output and refactor values."""
    import json
    with open(os.path.join(GOLDEN, "analysis_sha256.json")) as f:
        return json.load(f)


REFACT_SEED = 9000   # drifted value set of the golden refactor values (make_golden.py)


def order_sensitive_values(an, indices, indptr, data, span=32):
    """A value set near `data` on which the two FP64 product orders of the
    MC64 scaling, (a * rs) * csp (NICSLU's refactor) and a * (rs * csp) (its
    analysis), round to different FP32 values wherever that can happen.

    Entry k of the caller's CSC (column j, row i) is scaled by rs[j] and
    csp[col_perm_inv[i]].  Each entry is moved to a value whose scaled
    product sits at the midpoint of two adjacent FP32 numbers, then stepped
    one FP64 ulp at a time (up to `span` steps each way) until the two orders
    land on opposite sides of it.  Returns (values, mask of the entries where
    the orders differ in FP32).  Only entries whose scales are both not
    powers of two can differ at all."""
    n = len(indptr) - 1
    col = np.repeat(np.arange(n), np.diff(np.asarray(indptr, dtype=np.int64)))
    s = np.asarray(an.row_scale)[col]
    c = np.asarray(an.col_scale_perm)[np.asarray(an.col_perm_inv)[indices]]
    f = (data * s * c).astype(np.float32)
    mid = (f.astype(np.float64) + np.nextafter(f, np.float32(np.inf)).astype(np.float64)) / 2
    a = mid / (s * c)
    out = np.array(data, dtype=np.float64, copy=True)
    found = np.zeros(out.size, dtype=bool)
    lo, hi = a.copy(), a.copy()
    with np.errstate(all="ignore"):
        for _ in range(span):
            for cand in (lo, hi):
                o0 = ((cand * s) * c).astype(np.float32)
                o1 = (cand * (s * c)).astype(np.float32)
                m = (o0 != o1) & ~found & np.isfinite(cand) & (data != 0)
                out[m] = cand[m]
                found |= m
            lo = np.nextafter(lo, -np.inf)
            hi = np.nextafter(hi, np.inf)
    return out, found


def unsym_case(n=400, seed=31):
    """A random unsymmetric matrix with magnitudes spread over six decades:
    MC64's row and column scales are rarely powers of two on it."""
    rng = np.random.default_rng(seed)
    A = sp.random(n, n, density=0.02, random_state=int(rng.integers(1 << 30)), format="csc")
    A.data[:] = 10.0 ** rng.uniform(-3, 3, A.data.size) * rng.choice([-1.0, 1.0], A.data.size)
    A = A + sp.csc_matrix((10.0 ** rng.uniform(-1, 2, n), (rng.permutation(n), np.arange(n))), shape=(n, n))
    A = sp.csc_matrix(A)
    A.sum_duplicates()
    A.sort_indices()
    return A.data.copy(), A.indices.astype(np.int32), A.indptr.astype(np.int32), n


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
