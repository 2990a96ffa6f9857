"""Autotuner for the schedule parameters of a plan.

* ``fuse_units`` — level fusion: a factorization level needing at most this
  many lanes runs inside one workgroup (one barrier per level) instead of
  its own grid launch; solve levels fuse at 4 rows per thread.
* ``block_threads`` — threads of the fused workgroup and of the batched
  per-matrix workgroup (the lanes that share one level's items).
* ``lds_windows`` — how many runs of steps are factored with their values
  staged in LDS (counted from the last step; -1 = none).
* ``krylov`` (batched solve) — the FGMRES restart length of the refinement
  (-1 = plain refinement).

None of these changes a result: factors are bit-identical under every
schedule (only the order of independent work changes).

Each candidate plan is timed on the device with HIP events over the real
workload (single-matrix factorization, one same-pattern batched launch, or
one fused launch over an uneven group of patterns, whose workgroup size is
the tile every member runs on) and the fastest is returned.  Results of
every candidate are kept for reporting.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import torch

from .device import Plan

FUSE_CANDIDATES = (256, 512, 1024, 2048, 4096, 8192, 16384)
THREAD_CANDIDATES = (256, 512, 1024)


WINDOW_CANDIDATES = (0, -1, 1)
KRYLOV_CANDIDATES = (2, -1, 1, 4)
# uneven groups: members on the single-matrix path, against the automatic
# choice the other stages run with (one in a group of four or more)
SPLIT_CANDIDATES = (0, 2)


@dataclass
class TuneResult:
    fuse_units: int
    block_threads: int
    ms: float
    table: list = field(default_factory=list)
    lds_windows: int = 0
    krylov: int = 0
    split: int = -1   # uneven groups: members on the single-matrix path (-1: automatic)

    def plan(self, analysis):
        """A plan with the chosen parameters."""
        return Plan(analysis, fuse_units=self.fuse_units, block_threads=self.block_threads,
                    lds_windows=self.lds_windows, krylov=self.krylov)


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


def tune_factor(analysis, values=None, fuse=FUSE_CANDIDATES, threads=THREAD_CANDIDATES,
                windows=WINDOW_CANDIDATES, reps=5):
    """Best (fuse_units, block_threads, lds_windows) for the single-matrix
    factorization: every combination of fuse and threads at the default
    window count, then the window counts at the best of those."""
    table = []
    base = torch.from_numpy(analysis.val.copy()).cuda() if values is None else values
    best = None

    def measure(f, t, wnd):
        plan = Plan(analysis, fuse_units=f, block_threads=t, lds_windows=wnd)
        lu = base.clone()

        def run():
            lu.copy_(base)
            plan.factor(lu)

        ms = _time(run, reps)
        table.append((f, t, plan.segments, ms, wnd))
        plan.close()
        return ms

    for t in threads:
        for f in fuse:
            ms = measure(f, t, 0)
            if best is None or ms < best[2]:
                best = (f, t, ms, 0)
    for wnd in windows:
        if wnd == 0:
            continue
        ms = measure(best[0], best[1], wnd)
        if ms < best[2]:
            best = (best[0], best[1], ms, wnd)
    return TuneResult(best[0], best[1], best[2], table, lds_windows=best[3])


def tune_batch(analysis, ax, b, tol=1e-10, maxit=10, threads=THREAD_CANDIDATES, krylov=KRYLOV_CANDIDATES,
               reps=3):
    """Best (block_threads, krylov) for the batched launch (ax: (B, nnz_a),
    b: (B, n)): the thread counts at the default refinement, then the
    restart lengths at the best thread count."""
    table = []
    best = None

    def measure(t, m):
        plan = Plan(analysis, block_threads=t, krylov=m)
        ms = _time(lambda: plan.batch_solve(ax, b, tol=tol, maxit=maxit), reps, warmup=1)
        table.append((0, t, 0, ms, m))
        plan.close()
        return ms

    for t in threads:
        ms = measure(t, 0)
        if best is None or ms < best[2]:
            best = (0, t, ms, 0)
    for m in krylov:
        ms = measure(best[1], m)
        if ms < best[2]:
            best = (0, best[1], ms, m)
    return TuneResult(best[0], best[1], best[2], table, krylov=best[3])


def tune_group(analyses, axs, bs, tol=1e-10, maxit=10, threads=THREAD_CANDIDATES, krylov=KRYLOV_CANDIDATES,
               windows=WINDOW_CANDIDATES, splits=SPLIT_CANDIDATES, reps=3, plan_pool=None):
    """Best group-wide (block_threads, krylov, lds_windows, split) for one
    fused launch over distinct patterns (glu_group_solve): the launch's
    workgroup size is the tile every member runs on, and each candidate
    rebuilds all member plans.  Thread counts at the defaults first, then
    restart lengths, then window counts, each at the best so far, then how
    many of the heaviest members run on the single-matrix path beside the
    launch (results are the same bits for every split).  ``plan_pool`` (an
    executor) builds the member plans in parallel."""
    from .device import Group
    table = []
    best = None

    def measure(t, m, wnd, split=None):
        mk = lambda an: Plan(an, block_threads=t, krylov=m, lds_windows=wnd)
        plans = list(plan_pool.map(mk, analyses)) if plan_pool is not None else [mk(a) for a in analyses]
        g = Group(plans, split=split)
        ms = _time(lambda: g.solve(axs, bs, tol=tol, maxit=maxit), reps, warmup=1)
        table.append((0, t, 0, ms, (m, wnd, -1 if split is None else split)))
        g.close()
        for p in plans:
            p.close()
        return ms

    for t in threads:
        ms = measure(t, 0, 0)
        if best is None or ms < best[2]:
            best = (0, t, ms, 0, 0)
    for m in krylov:
        ms = measure(best[1], m, 0)
        if ms < best[2]:
            best = (0, best[1], ms, m, 0)
    for wnd in windows:
        if wnd == 0:
            continue
        ms = measure(best[1], best[3], wnd)
        if ms < best[2]:
            best = (0, best[1], ms, best[3], wnd)
    split = -1
    for sp in splits:
        ms = measure(best[1], best[3], best[4], sp)
        if ms < best[2]:
            best = (0, best[1], ms, best[3], best[4])
            split = sp
    return TuneResult(best[0], best[1], best[2], table, lds_windows=best[4], krylov=best[3], split=split)


def report(res: TuneResult) -> str:
    lines = [f"{'fuse_units':>10} {'threads':>8} {'segments':>9} {'ms':>10} {'extra':>6}"]
    best = min(res.table, key=lambda r: r[3]) if res.table else None
    for row in res.table:
        f, t, s, ms, x = row
        mark = " *" if row is best else ""
        lines.append(f"{f:>10} {t:>8} {s:>9} {ms:>10.4f} {str(x):>6}{mark}")
    return "\n".join(lines)


__all__ = ["tune_factor", "tune_batch", "tune_group", "report", "TuneResult", "np"]
