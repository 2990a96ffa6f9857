"""Seeded generator of modified-nodal-analysis (MNA) circuit matrices.

The reference ships one example input (``src/add32.mtx``, n=4960) and no
generator; every synthetic workload used by the tests and the benchmark
comes from this module so that shapes and seeds are reproducible.

Structure of a generated matrix ``A`` (n = n_nodes + n_vsrc):

* node graph: a spanning chain ``i -- i+1`` (connectivity) plus a
  locality-constrained configuration-model pairing.  Each node draws
  ``extra`` stubs; stubs are sorted by ``node + U[0, window)`` and adjacent
  stubs are paired, so branches join nodes that are close in index.  A
  globally random pairing would be an expander graph whose LU fill is
  dense; circuits are local, and so is this graph.
* rails: ``n_rails`` randomly placed nodes joined to ``rail_deg`` random
  nodes each (supply nets, the high-degree columns of circuit matrices).
* every branch is a conductance ``g ~ logU[g_lo, g_hi]`` stamped
  symmetrically (``+g`` on both diagonals, ``-g`` off-diagonal); every node
  gets a shunt ``logU[g_lo, g_hi]``, so the node block is strictly
  diagonally dominant.
* ``n_vsrc`` voltage sources at distinct nodes ``p``:
  ``A[p, nn+s] = +1`` and ``A[nn+s, p] = sigma_s`` with ``sigma_s = +-1``.
  The source rows have a structurally zero diagonal, which is what makes
  the MC64 matching step of the host preprocess necessary.

The pattern is structurally symmetric; values are not (the source stamps).
Output is CSC with summed duplicates and sorted row indices, the form the
reference's Python front end produces (``pyglu/__init__.py:30-35``).
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np


@dataclass(frozen=True)
class MNASpec:
    n_nodes: int
    n_vsrc: int
    extra_lo: int = 0        # extra stubs per node on top of the chain's 2
    extra_hi: int = 4
    n_rails: int = 3
    rail_deg: int = 300
    window: float = 24.0
    g_lo: float = 1e-3
    g_hi: float = 1e3

    @property
    def n(self) -> int:
        return self.n_nodes + self.n_vsrc


def _log_uniform(rng, lo, hi, size):
    return np.exp(rng.uniform(math.log(lo), math.log(hi), size))


def mna_csc(spec: MNASpec, seed: int):
    """Return ``(data f64, indices i32, indptr i32, n)`` of one MNA matrix."""
    rng = np.random.default_rng(seed)
    nn = spec.n_nodes
    if nn < 2:
        raise ValueError("need at least two circuit nodes")
    if spec.n_vsrc > nn:
        raise ValueError("more voltage sources than nodes")

    # spanning chain
    ea = [np.arange(nn - 1, dtype=np.int64)]
    eb = [np.arange(1, nn, dtype=np.int64)]

    # locality-constrained configuration-model pairing
    extra = rng.integers(spec.extra_lo, spec.extra_hi + 1, nn)
    stubs = np.repeat(np.arange(nn, dtype=np.int64), extra)
    if stubs.size >= 2:
        key = stubs + rng.uniform(0.0, spec.window, stubs.size)
        s = stubs[np.argsort(key, kind="stable")]
        if s.size % 2:
            s = s[:-1]
        ea.append(s[0::2])
        eb.append(s[1::2])

    # rails
    if spec.n_rails > 0:
        rails = rng.choice(nn, size=min(spec.n_rails, nn), replace=False)
        for r in rails:
            k = min(spec.rail_deg, nn - 1)
            nb = rng.choice(nn - 1, size=k, replace=False)
            nb = nb + (nb >= r)          # skip the rail itself
            ea.append(np.full(k, r, dtype=np.int64))
            eb.append(nb.astype(np.int64))

    a = np.concatenate(ea)
    b = np.concatenate(eb)
    keep = a != b
    a, b = a[keep], b[keep]
    lo = np.minimum(a, b)
    hi = np.maximum(a, b)
    code = np.unique(lo * nn + hi)
    lo = code // nn
    hi = code % nn
    g = _log_uniform(rng, spec.g_lo, spec.g_hi, code.size)

    shunt = _log_uniform(rng, spec.g_lo, spec.g_hi, nn)
    diag = shunt.copy()
    np.add.at(diag, lo, g)
    np.add.at(diag, hi, g)

    rows = [np.arange(nn), lo, hi]
    cols = [np.arange(nn), hi, lo]
    vals = [diag, -g, -g]

    nv = spec.n_vsrc
    if nv > 0:
        p = rng.choice(nn, size=nv, replace=False)
        sigma = np.where(rng.random(nv) < 0.5, -1.0, 1.0)
        s = nn + np.arange(nv)
        rows += [p, s]
        cols += [s, p]
        vals += [np.ones(nv), sigma]

    r = np.concatenate(rows).astype(np.int64)
    c = np.concatenate(cols).astype(np.int64)
    v = np.concatenate(vals).astype(np.float64)
    n = spec.n

    # CSC assembly: sort by (col, row); the (row, col) pairs are unique by
    # construction (unique edges, one diagonal per node, distinct sources).
    order = np.lexsort((r, c))
    r, c, v = r[order], c[order], v[order]
    indptr = np.zeros(n + 1, dtype=np.int64)
    np.add.at(indptr, c + 1, 1)
    indptr = np.cumsum(indptr)
    return (v, r.astype(np.int32), indptr.astype(np.int32), n)


# Workload shapes (sizes, degree mixes, rail counts and seeds).
SPEC_SMALL = MNASpec(n_nodes=4920, n_vsrc=40, extra_lo=0, extra_hi=3,
                     n_rails=3, rail_deg=300)
SPEC_LARGE = MNASpec(n_nodes=496000, n_vsrc=4000, extra_lo=0, extra_hi=7,
                     n_rails=50, rail_deg=5000)
SPEC_BATCH = MNASpec(n_nodes=19840, n_vsrc=160, extra_lo=0, extra_hi=5,
                     n_rails=5, rail_deg=1000)
SPEC_GRAD = MNASpec(n_nodes=49600, n_vsrc=400, extra_lo=0, extra_hi=5,
                    n_rails=10, rail_deg=2000)

SEED_SMALL, SEED_LARGE, SEED_BATCH, SEED_GRAD = 1, 2, 3, 5


def uneven_specs(count: int = 64, seed0: int = 4000):
    """Specs of the uneven batch: n log-uniform in [1k, 50k], avg degree 4..8."""
    out = []
    for m in range(count):
        rng = np.random.default_rng(seed0 + m)
        n = int(10 ** rng.uniform(3.0, math.log10(5e4)))
        deg = int(rng.integers(4, 9))
        nv = max(1, n // 125)
        nn = n - nv
        # chain gives 2; pairing adds `deg - 2` on average
        extra_hi = max(0, 2 * (deg - 2))
        out.append((MNASpec(n_nodes=nn, n_vsrc=nv, extra_lo=0, extra_hi=extra_hi,
                            n_rails=3, rail_deg=max(8, n // 40)), seed0 + m))
    return out


def perturb_values(data: np.ndarray, seed: int, rel: float = 0.1) -> np.ndarray:
    """Newton-step drift: ``data * (1 + U[-rel, rel])`` per stored entry."""
    rng = np.random.default_rng(seed)
    return data * (1.0 + rng.uniform(-rel, rel, data.size))


def rhs(n: int, seed: int) -> np.ndarray:
    return np.random.default_rng(seed).uniform(-1.0, 1.0, n)
