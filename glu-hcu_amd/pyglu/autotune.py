"""Autotuner for the two schedule parameters of a plan.

* ``fuse_units`` — level fusion: a factorization level needing at most this
  many lanes runs inside one workgroup (one barrier per level) instead of
  its own grid launch; solve levels fuse at 4 rows per thread.
* ``block_threads`` — threads of the fused workgroup and of the batched
  per-matrix workgroup (the lanes that share one level's items).

Each candidate plan is timed on the device with HIP events over the real
workload (single-matrix factorization, or one batched launch) and the
fastest is returned.  Results of every candidate are kept for reporting.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import torch

from .device import Plan

FUSE_CANDIDATES = (256, 512, 1024, 2048, 4096, 8192, 16384)
THREAD_CANDIDATES = (256, 512, 1024)


@dataclass
class TuneResult:
    fuse_units: int
    block_threads: int
    ms: float
    table: list = field(default_factory=list)


def _time(fn, reps=5, warmup=2):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(reps):
        fn()
    e.record()
    torch.cuda.synchronize()
    return s.elapsed_time(e) / reps


def tune_factor(analysis, values=None, fuse=FUSE_CANDIDATES, threads=THREAD_CANDIDATES, reps=5):
    """Best (fuse_units, block_threads) for the single-matrix factorization."""
    table = []
    base = torch.from_numpy(analysis.val.copy()).cuda() if values is None else values
    best = None
    for t in threads:
        for f in fuse:
            plan = Plan(analysis, fuse_units=f, block_threads=t)
            lu = base.clone()

            def run():
                lu.copy_(base)
                plan.factor(lu)

            ms = _time(run, reps)
            table.append((f, t, plan.segments, ms))
            if best is None or ms < best[2]:
                best = (f, t, ms)
            plan.close()
    return TuneResult(best[0], best[1], best[2], table)


def tune_batch(analysis, ax, b, tol=1e-10, maxit=10, threads=THREAD_CANDIDATES, reps=3):
    """Best block_threads for the batched launch (ax: (B, nnz_a), b: (B, n))."""
    table = []
    best = None
    for t in threads:
        plan = Plan(analysis, block_threads=t)
        ms = _time(lambda: plan.batch_solve(ax, b, tol=tol, maxit=maxit), reps, warmup=1)
        table.append((0, t, 0, ms))
        if best is None or ms < best[2]:
            best = (0, t, ms)
        plan.close()
    return TuneResult(best[0], best[1], best[2], table)


def report(res: TuneResult) -> str:
    lines = [f"{'fuse_units':>10} {'threads':>8} {'segments':>9} {'ms':>10}"]
    for f, t, s, ms in res.table:
        mark = " *" if (f, t) == (res.fuse_units, res.block_threads) else ""
        lines.append(f"{f:>10} {t:>8} {s:>9} {ms:>10.4f}{mark}")
    return "\n".join(lines)


__all__ = ["tune_factor", "tune_batch", "report", "TuneResult", "np"]
