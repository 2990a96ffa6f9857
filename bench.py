#!/usr/bin/env python
"""Benchmark: batched same-pattern refactorization + FP64 refinement.

Workload: 256 MNA matrices sharing one n=20,000 (~112k nnz) pattern, values
= base * (1 + U[-0.1, 0.1]) per entry, one RHS each; a step is one batched
solve: refresh the values into the LU pattern (frozen MC64 scales and
permutation), factor in FP32 (the wide leading steps as grid launches over
an interleaved copy of the batch, 16 matrices per 64-byte line, the rest in
one workgroup per matrix), then solve in FP32 and refine in FP64 (FGMRES
preconditioned by the FP32 factors, one workgroup per matrix) until the
normwise backward error is <= 1e-10.  Inputs are resident on the GPU before
the timed region.

Reported beside it (same JSON line):
* the rate of looping the plain FP32 drop-in path (single-matrix
  level-scheduled factor + refinement) over the same matrices;
* the batched refactor alone (refresh + factor, no solve);
* the batch rate including the host-to-device copy of values and RHS;
* the plain FP32 factor + solve of the n=4,960 MNA matrix;
* the single-matrix numeric refactor time of the n=500,000 matrix;
* the uneven group (64 distinct patterns, n in [1k, 50k]) in one launch;
* forward + backward of the differentiable solve at n=50,000;
* the batch step's HBM roofline fraction;
* the reference NICSLU CPU refactor rate on the host cores.

python bench.py [--gpus 1] [--steps K] [--warmup W]

--copy-probe adds one device copy of 256 MiB (read) + 256 MiB (write) of
known size, whose counter readings calibrate FETCH_SIZE / WRITE_SIZE.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "glu-hcu_amd"), ROOT]

HBM_PEAK_GBS = 8000.0


def alg_bytes_factor(nnz_lu, n, share):
    """Compulsory bytes of one FP32 factorization: read + write L\\U values,
    plus the pattern (CSC + CSR + diagonal pointers + level list) amortised
    over the matrices that share it."""
    pattern = 4 * (2 * nnz_lu + 4 * n + 2) + 4 * n
    return 8 * nnz_lu + pattern / share


def alg_bytes_refine(nnz_a, nnz_lu, n, share):
    """One refinement iteration: FP64 SpMV over A (values + column index),
    two FP32 triangular solves over the factors, FP64 vectors."""
    return 12 * nnz_a + 8 * (n + 1) / share + 8 * nnz_lu + 56 * n


def fp64_solution(an, data, indices, indptr, b):
    """x* for A x = b on the host: an FP64 sparse LU (SuperLU) of A with its
    rows and columns in the analysis's MC64 + AMD order (natural order, no
    pivoting: the permuted matrix carries the matching on its diagonal),
    then two refinement steps with the residual in the host's extended
    precision, so x* is accurate far below the errors it measures.
    Independent of the GPU path: SuperLU factors A's FP64 values itself."""
    import numpy as np
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla
    n = an.n
    A = sp.csc_matrix((data, indices, indptr), shape=(n, n))
    rp = np.asarray(an.col_perm, dtype=np.int64)   # GLU row r = caller row col_perm[r]
    cq = np.asarray(an.row_perm, dtype=np.int64)   # GLU column c = caller column row_perm[c]
    lu = sla.splu(sp.csc_matrix(A[rp, :][:, cq]), permc_spec="NATURAL", diag_pivot_thresh=0.0,
                  options=dict(Equil=False))

    def solve(r):
        x = np.empty(n)
        x[cq] = lu.solve(r[rp])
        return x

    b64 = np.asarray(b, dtype=np.float64)
    xs = solve(b64)
    Al = sp.csr_matrix(A).astype(np.longdouble)
    for _ in range(2):
        r = b64.astype(np.longdouble) - Al @ xs.astype(np.longdouble)
        xs = xs + solve(r.astype(np.float64))
    return xs


def accuracy(an, data, indices, indptr, b, x):
    """Forward error |x - x*|_inf / |x*|_inf against fp64_solution, the
    normwise backward error eta = |r|_inf / (|A|_inf |x|_inf + |b|_inf) and
    the componentwise backward error omega = max_i |r_i| / (|A||x| + |b|)_i,
    with r = b - A x formed in extended precision."""
    import numpy as np
    import scipy.sparse as sp
    n = an.n
    x = np.asarray(x, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    xs = fp64_solution(an, data, indices, indptr, b)
    A = sp.csr_matrix(sp.csc_matrix((data, indices, indptr), shape=(n, n)))
    r = np.abs((b.astype(np.longdouble) - A.astype(np.longdouble) @ x.astype(np.longdouble)).astype(np.float64))
    absA = abs(A)
    den = absA @ np.abs(x) + np.abs(b)
    with np.errstate(divide="ignore", invalid="ignore"):
        comp = np.where(den > 0, r / den, np.where(r > 0, np.inf, 0.0))
    return {"fwd": float(np.abs(x - xs).max() / np.abs(xs).max()),
            "eta": float(r.max() / (absA.sum(axis=1).max() * np.abs(x).max() + np.abs(b).max())),
            "omega": float(comp.max())}


def summarize(accs):
    """max / median of the accuracy figures over a set of results."""
    import numpy as np
    out = {}
    for k in ("fwd", "eta", "omega"):
        v = np.array([a[k] for a in accs])
        out[f"{k}_max"] = float(v.max())
        out[f"{k}_median"] = float(np.median(v))
    out["count"] = len(accs)
    return out


def fwd_error(an, data, indices, indptr, b, x):
    """Forward error |x - x*|_inf / |x*|_inf against fp64_solution."""
    import numpy as np
    xs = fp64_solution(an, data, indices, indptr, b)
    return float(np.abs(np.asarray(x, dtype=np.float64) - xs).max() / np.abs(xs).max())


def _events(fn, reps):
    """Median time (ms) of fn() over reps, HIP events on the current stream."""
    import torch
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
    for s_, e_ in evs:
        s_.record()
        fn()
        e_.record()
    torch.cuda.synchronize()
    return sorted(s_.elapsed_time(e_) for s_, e_ in evs)[reps // 2]


def bench_small(dev, log):
    """Plain GLU FP32 factor + FP32 solve of the n=4960 MNA matrix."""
    import torch
    from pyglu import mna
    from pyglu.analysis import Analysis
    from pyglu.device import Plan
    d, i, p, n = mna.mna_csc(mna.SPEC_SMALL, mna.SEED_SMALL)
    an = Analysis(d, i, p, n)
    plan = Plan(an)
    ax = torch.from_numpy(d).to(dev)
    base = plan.refresh(ax).reshape(-1).clone()
    lu = base.clone()
    b32 = torch.from_numpy(mna.rhs(n, 101)).to(dev).to(torch.float32)
    x32 = torch.empty_like(b32)

    def step():
        lu.copy_(base)
        plan.factor(lu)
        plan.solve(lu, b32, x32)

    for _ in range(3):
        step()
    ms = _events(step, 11)
    fms = _events(lambda: (lu.copy_(base), plan.factor(lu)), 11)
    b64 = mna.rhs(n, 101)
    bd = torch.from_numpy(b64).to(dev)
    xr, it, eta = plan.refine(ax, lu, bd, tol=1e-10, maxit=20)
    xt, itt, etat = plan.refine(ax, lu, bd, tol=1e-13, maxit=30)
    acc32 = accuracy(an, d, i, p, b64, x32.cpu().numpy())
    accr = accuracy(an, d, i, p, b64, xr.cpu().numpy())
    acct = accuracy(an, d, i, p, b64, xt.cpu().numpy())
    out = {"n": n, "nnz": int(d.size), "nnz_lu": an.nnz_lu, "levels": an.num_lev,
           "factor_solve_ms": round(ms, 4), "factor_ms": round(fms, 4),
           "us_per_level": round(fms * 1e3 / max(an.num_lev, 1), 2),
           "accuracy_fp32_solve": acc32,
           "accuracy_refined_1e-10": {**accr, "solves": it},
           "accuracy_refined_1e-13": {**acct, "solves": itt}}
    log(f"[bench] n=4960: factor+solve {ms:.3f} ms (factor {fms:.3f} ms, {an.num_lev} levels); forward error "
        f"vs FP64: FP32 solve {acc32['fwd']:.1e}, refined to 1e-10 {accr['fwd']:.1e} ({it} solves), "
        f"to 1e-13 {acct['fwd']:.1e} ({itt} solves)")
    plan.close()
    return out


def bench_uneven(dev, log, tol, maxit):
    """64 distinct MNA patterns (n log-uniform in [1k, 50k]) in one launch."""
    import numpy as np
    import torch
    from pyglu import mna
    from pyglu.analysis import analyze_many
    from pyglu.device import Group, Plan
    mats = [mna.mna_csc(spec, seed) for spec, seed in mna.uneven_specs()]
    threads = min(16, os.cpu_count() or 1)
    t0 = time.perf_counter()
    ans = analyze_many(mats, threads=threads)   # host analysis, one pattern per thread
    t_an = time.perf_counter() - t0
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(threads) as ex:   # plan tables are built on the host, one pattern per thread
        plans = list(ex.map(Plan, ans))
    axs, bs = [], []
    for m, (d, i, p, n) in enumerate(mats):
        axs.append(torch.from_numpy(d).to(dev))
        bs.append(torch.from_numpy(mna.rhs(n, 4500 + m)).to(dev))
    g = Group(plans)
    t_plan = time.perf_counter() - t0 - t_an

    def launch_ms(grp, reps=5):
        grp.solve(axs, bs, tol=tol, maxit=maxit)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            res = grp.solve(axs, bs, tol=tol, maxit=maxit)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps * 1e3, res

    default_ms, _ = launch_ms(g)
    g0 = Group(plans, split=0)   # every member in the fused launch
    fused_only_ms, _ = launch_ms(g0)
    g0.close()
    # group-wide tiles (workgroup size), refinement restart and LDS windows,
    # autotuned on this group; the tuned group is the one measured
    from pyglu import autotune
    t0 = time.perf_counter()
    with ThreadPoolExecutor(threads) as ex:
        tune = autotune.tune_group(ans, axs, bs, tol=tol, maxit=maxit, plan_pool=ex)
        tuned_plans = list(ex.map(lambda a: Plan(a, block_threads=tune.block_threads, krylov=tune.krylov,
                                                lds_windows=tune.lds_windows), ans))
    t_tune = time.perf_counter() - t0
    gt = Group(tuned_plans, split=None if tune.split < 0 else tune.split)
    ms, (xs, _, iters, eta) = launch_ms(gt)
    it_h, eta_h = iters.cpu().numpy(), eta.cpu().numpy()
    accs = [accuracy(ans[m], d, i, p, mna.rhs(n, 4500 + m), xs[m].cpu().numpy())
            for m, (d, i, p, n) in enumerate(mats)]
    # the plain loop over the same matrices
    t0 = time.perf_counter()
    for p, ax, b in zip(plans, axs, bs):
        lu = p.refresh(ax).reshape(-1)
        p.factor(lu)
        p.refine(ax, lu, b, tol=tol, maxit=maxit)
    torch.cuda.synchronize()
    loop_ms = (time.perf_counter() - t0) * 1e3
    out = {"count": len(plans), "sum_n": int(sum(p.n for p in plans)), "max_n": int(max(p.n for p in plans)),
           "max_levels": int(max(p.num_lev for p in plans)), "launch_ms": round(ms, 3),
           "refactors_per_s": round(len(plans) / ms * 1e3, 1), "plain_loop_ms": round(loop_ms, 3),
           "default_params_launch_ms": round(default_ms, 3), "fused_only_launch_ms": round(fused_only_ms, 3),
           "autotuned": {"block_threads": tune.block_threads, "krylov": tune.krylov,
                         "lds_windows": tune.lds_windows, "split": tune.split, "candidates": len(tune.table),
                         "gain_vs_default": round(default_ms / ms, 3), "tune_s": round(t_tune, 2)},
           "refine_iters_max": int(it_h.max()), "eta_max": float(eta_h.max()),
           "accuracy": summarize(accs),
           "correct": bool(np.all(eta_h <= tol)), "host_analysis_s": round(t_an, 2),
           "host_analysis_threads": threads, "plans_s": round(t_plan, 2)}
    log(f"[bench] uneven group: {ms:.3f} ms per launch autotuned (threads {tune.block_threads}, krylov "
        f"{tune.krylov}, windows {tune.lds_windows}, split {tune.split}; defaults {default_ms:.3f} ms, "
        f"all members in the fused launch {fused_only_ms:.3f} ms; "
        f"plain loop {loop_ms:.1f} ms), "
        f"iters max {it_h.max()}, eta max {eta_h.max():.2e}, forward error max {out['accuracy']['fwd_max']:.1e}")
    g.close()
    gt.close()
    for p in plans + tuned_plans:
        p.close()
    return out


def bench_autograd(dev, log):
    """Forward + backward of spsolve_autograd at n=50,000."""
    import numpy as np
    import torch
    from pyglu import mna
    from pyglu.analysis import Analysis
    from pyglu.device import Plan, spsolve_autograd
    d, i, p, n = mna.mna_csc(mna.SPEC_GRAD, mna.SEED_GRAD)
    plan = Plan(Analysis(d, i, p, n))
    xstar = torch.from_numpy(np.random.default_rng(5000).standard_normal(n)).to(dev)
    b0 = torch.from_numpy(mna.rhs(n, 5001)).to(dev)
    vals0 = torch.from_numpy(d).to(dev)

    def step():
        v = vals0.clone().requires_grad_(True)
        b = b0.clone().requires_grad_(True)
        x = spsolve_autograd(v, b, plan)
        loss = ((x - xstar) ** 2).sum()
        loss.backward()
        return v.grad, b.grad

    step()
    torch.cuda.synchronize()
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        step()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / reps * 1e3
    log(f"[bench] autograd n={n}: forward+backward {ms:.3f} ms")
    plan.close()
    return {"n": n, "nnz": int(d.size), "forward_backward_ms": round(ms, 3)}


def bench_large(dev, log, tol):
    """n=500k MNA matrix: single refactor (median of 5, HIP events) and the
    FP64 refinement to tol (wall clock of one glu_refine call)."""
    import torch
    from pyglu import mna
    from pyglu.analysis import Analysis
    from pyglu.device import Plan
    t0 = time.perf_counter()
    dL, iL, pL, nL = mna.mna_csc(mna.SPEC_LARGE, mna.SEED_LARGE)
    anL = Analysis(dL, iL, pL, nL)
    planL = Plan(anL)
    t_an = time.perf_counter() - t0
    axL = torch.from_numpy(dL).to(dev)
    luL = planL.refresh(axL).reshape(-1)
    base = luL.clone()
    planL.factor(luL)
    torch.cuda.synchronize()
    reps = 5
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
    for s_, e_ in evs:
        luL.copy_(base)
        s_.record()
        planL.factor(luL)
        e_.record()
    torch.cuda.synchronize()
    ms = sorted(s_.elapsed_time(e_) for s_, e_ in evs)[reps // 2]
    bL = torch.from_numpy(mna.rhs(nL, 102)).to(dev)
    planL.refine(axL, luL, bL, tol=tol, maxit=20)   # warm (scratch allocation)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    xL, itL, etaL = planL.refine(axL, luL, bL, tol=tol, maxit=20)
    refine_ms = (time.perf_counter() - t0) * 1e3
    acc = accuracy(anL, dL, iL, pL, mna.rhs(nL, 102), xL.cpu().numpy())
    log(f"[bench] n=500k: analysis+plan {t_an:.1f}s refactor {ms:.3f} ms, refine {itL} iters {refine_ms:.1f} ms "
        f"eta {etaL:.2e}, forward error {acc['fwd']:.1e}, componentwise backward error {acc['omega']:.1e}")
    out = {"n": nL, "nnz": int(dL.size), "nnz_lu": anL.nnz_lu, "levels": anL.num_lev,
           "lds_windows": planL.lds_window_count, "refactor_ms": round(ms, 3), "refine_iters": itL, "eta": etaL,
           "refine_ms": round(refine_ms, 3), "accuracy": acc, "host_analysis_s": round(t_an, 2)}
    planL.close()
    return out


def cpu_cores():
    """(threads to use, affinity cores, cgroup CPU quota or None, online
    cores): the cores this process may run on (its affinity mask), capped by
    a cgroup CPU quota when one is set, so that NICSLU's threads are not
    oversubscribed."""
    aff = len(os.sched_getaffinity(0))
    quota = None
    for path, parse in (("/sys/fs/cgroup/cpu.max", lambda s: s.split()),
                        ("/sys/fs/cgroup/cpu/cpu.cfs_quota_us", lambda s: [s.strip(), None])):
        try:
            with open(path) as f:
                q, per = parse(f.read())
            if per is None:
                with open("/sys/fs/cgroup/cpu/cpu.cfs_period_us") as f:
                    per = f.read().strip()
            if q not in ("max", "-1"):
                quota = max(1, int(int(q) // int(per)))
            break
        except (OSError, ValueError):
            continue
    return (min(aff, quota) if quota else aff), aff, quota, os.cpu_count()


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(ref, data, indices, indptr, n, ax_h, args, log):
    """NICSLU (the reference's vendored CPU solver, compiled from its sources
    into oracle/_ref) on the host cores, per src/nicslu/demo/demop_time.c:
    Analyze, Factorize (one thread, partial pivoting), then with T threads
    ReFactorize_MT per value set and one Solve; wall clock (CLOCK_MONOTONIC).
    The reported value is the batch configuration's refactor + solve rate
    (passes over the 256 value sets until about cpu_seconds of refactor +
    solve time); the n=4,960 matrix, the 64-pattern uneven group and, with
    --cpu-large, the n=500k matrix are timed beside it."""
    import numpy as np
    from pyglu import mna
    T, aff, quota, online = cpu_cores()
    B = ax_h.shape[0]
    t0 = time.perf_counter()
    spent, done, passes, first = 0.0, 0, 0, None
    while spent < args.cpu_seconds and passes < 1000:
        res = ref.nicslu_time(data, indices, indptr, n, ax_h, threads=T)
        first = first or res
        spent += B * (res["refactor_s"] + res["solve_s"])
        done += B
        passes += 1
    rate = done / spent
    keep = lambda r: {k: round(float(v) * 1e3, 4) for k, v in r.items() if k != "x"}
    configs = {"batch_n20k": {**{f"{k[:-2]}_ms": v for k, v in keep(first).items()},
                              "refactor_solve_per_s": round(rate, 2)}}
    d, i, p, ns = mna.mna_csc(mna.SPEC_SMALL, mna.SEED_SMALL)
    r = ref.nicslu_time(d, i, p, ns, np.stack([d] * 20), threads=T)
    configs["small_n4960"] = {f"{k[:-2]}_ms": v for k, v in keep(r).items()}
    tot = {"analyze_s": 0.0, "factorize_s": 0.0, "refactor_s": 0.0, "solve_s": 0.0}
    for m, (spec, seed) in enumerate(mna.uneven_specs()):
        d, i, p, nm = mna.mna_csc(spec, seed)
        r = ref.nicslu_time(d, i, p, nm, d[None, :], threads=T)
        for k in tot:
            tot[k] += float(r[k])
    configs["uneven_64"] = {f"{k[:-2]}_ms_sum": round(v * 1e3, 3) for k, v in tot.items()}
    if args.cpu_large:
        d, i, p, nl = mna.mna_csc(mna.SPEC_LARGE, mna.SEED_LARGE)
        r = ref.nicslu_time(d, i, p, nl, d[None, :], threads=T)
        configs["large_n500k"] = {f"{k[:-2]}_ms": v for k, v in keep(r).items()}
    log(f"[bench] NICSLU CPU: {rate:.1f} refactors/s on {T} threads ({done} refactors); per config {configs}")
    return {"value": round(rate, 2), "unit": "refactors/s", "cores": T, "kind": "reference",
            "sample": f"NICSLU (src/nicslu, -O2 -DSSE2 -DNO_ATOMIC) ReFactorize_MT + Solve, FP64, the {B} batch "
                      f"matrices x {passes} passes ({done} refactors, {spent:.1f} s of refactor + solve), "
                      f"{T} threads, {time.perf_counter() - t0:.1f} s wall incl. the other configurations",
            "cpu_model": cpu_model(), "affinity_cores": aff, "cgroup_quota_cpus": quota, "online_cpus": online,
            "configs": configs}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--maxit", type=int, default=10)
    ap.add_argument("--loop-sample", type=int, default=32,
                    help="matrices of the batch timed through the plain single-matrix loop")
    ap.add_argument("--no-large", action="store_true", help="skip the n=500k refactor timing")
    ap.add_argument("--no-cpu", action="store_true", help="skip the NICSLU CPU baseline")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--cpu-large", action="store_true",
                    help="also time NICSLU on the n=500k matrix (its Analyze alone takes minutes)")
    ap.add_argument("--block-threads", type=int, default=0)
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the n=4960, uneven-group and autograd side measurements")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run one 256 MiB device copy (counter calibration)")
    ap.add_argument("--single", choices=("small", "large"),
                    help="only the n=4960 factor + solve or the n=500k refactor + refinement measurement "
                         "(its JSON on stdout; for per-kernel profiles)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              f"(matrices shard across GPUs with no exchange, see the design notes)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from pyglu import mna
    from pyglu.analysis import Analysis
    from pyglu.device import Plan

    if not torch.cuda.is_available():
        print("bench.py needs a HIP device", file=sys.stderr)
        sys.exit(1)
    torch.cuda.set_device(0)
    dev = torch.device("cuda:0")
    log = lambda *a: print(*a, file=sys.stderr, flush=True)
    if args.single:
        res = bench_small(dev, log) if args.single == "small" else bench_large(dev, log, args.tol)
        print(json.dumps({"single": args.single, **res}))
        return

    # ---------------- workload ----------------------------------------
    B = args.batch
    data, indices, indptr, n = mna.mna_csc(mna.SPEC_BATCH, mna.SEED_BATCH)
    t0 = time.perf_counter()
    an = Analysis(data, indices, indptr, n)
    plan = Plan(an, block_threads=args.block_threads)
    log(f"[bench] n={n} nnz(A)={data.size} nnz(L+U)={an.nnz_lu} levels={an.num_lev} "
        f"updates={an.updates} analysis+plan {time.perf_counter() - t0:.2f}s")
    ax_h = np.stack([mna.perturb_values(data, 3000 + b) for b in range(B)])
    b_h = np.stack([mna.rhs(n, 3500 + b) for b in range(B)])
    ax = torch.from_numpy(ax_h).to(dev)
    bb = torch.from_numpy(b_h).to(dev)
    x = torch.empty(B, n, dtype=torch.float64, device=dev)
    lu = torch.empty(B, an.nnz_lu, dtype=torch.float32, device=dev)

    for _ in range(args.warmup):
        plan.batch_solve(ax, bb, tol=args.tol, maxit=args.maxit, lu=lu, x=x)
    torch.cuda.synchronize()

    K = args.steps
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(K)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(K)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(K):
        starts[k].record()
        _, _, iters, eta = plan.batch_solve(ax, bb, tol=args.tol, maxit=args.maxit, lu=lu, x=x)
        ends[k].record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = float(np.mean([s.elapsed_time(e) for s, e in zip(starts, ends)]))
    iters_h = iters.cpu().numpy()
    eta_h = eta.cpu().numpy()
    x_h = x.cpu().numpy()
    ok = bool(np.all(eta_h <= args.tol) and np.all(np.isfinite(x_h)))
    # every matrix's result on the host: forward error against an FP64
    # sparse LU, normwise and componentwise backward errors
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(min(16, os.cpu_count() or 1)) as ex:
        accs = list(ex.map(lambda b: accuracy(an, ax_h[b], indices, indptr, b_h[b], x_h[b]), range(B)))
    batch_acc = summarize(accs)
    ok = bool(ok and batch_acc["eta_max"] <= 2 * args.tol)
    log(f"[bench] batch accuracy over all {B} matrices: forward error max {batch_acc['fwd_max']:.2e} "
        f"(median {batch_acc['fwd_median']:.2e}), eta max {batch_acc['eta_max']:.2e}, componentwise max "
        f"{batch_acc['omega_max']:.2e}")
    value = B * K / wall
    log(f"[bench] batch: {wall / K * 1e3:.3f} ms/step (events {kernel_ms:.3f} ms) -> {value:.1f} refactors/s; "
        f"iters {iters_h.min()}..{iters_h.max()} max eta {eta_h.max():.2e} ok={ok}")

    if args.copy_probe:
        src = torch.empty(64 << 20, dtype=torch.float32, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"[bench] copy probe: {src.numel() * 4} bytes read, {dst.numel() * 4} bytes written")
        del src, dst

    # the same step with the value sets and RHS crossing PCIe first
    ax_pin = torch.from_numpy(ax_h).pin_memory()
    b_pin = torch.from_numpy(b_h).pin_memory()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(K):
        ax.copy_(ax_pin, non_blocking=True)
        bb.copy_(b_pin, non_blocking=True)
        plan.batch_solve(ax, bb, tol=args.tol, maxit=args.maxit, lu=lu, x=x)
    torch.cuda.synchronize()
    pcie_rate = B * K / (time.perf_counter() - t0)
    log(f"[bench] batch incl. H2D of values + rhs: {pcie_rate:.1f} refactors/s")
    del ax_pin, b_pin

    # ---------------- plain FP32 drop-in loop ---------------------------
    S = max(1, min(args.loop_sample, B))
    one_lu = torch.empty(an.nnz_lu, dtype=torch.float32, device=dev)
    one_x = torch.empty(n, dtype=torch.float64, device=dev)

    def single(b):
        plan.refresh(ax[b], one_lu)
        plan.factor(one_lu)
        return plan.refine(ax[b], one_lu, bb[b], x=one_x, tol=args.tol, maxit=args.maxit)

    for b in range(min(2, B)):
        single(b)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    loop_eta = 0.0
    for b in range(S):
        _, _, e = single(b)
        loop_eta = max(loop_eta, e)
    torch.cuda.synchronize()
    loop_rate = S / (time.perf_counter() - t0)
    # factor-only time of the single-matrix path (events)
    fs, fe = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    plan.refresh(ax[0], one_lu)
    fs.record()
    for _ in range(5):
        plan.factor(one_lu)
    fe.record()
    torch.cuda.synchronize()
    single_factor_ms = fs.elapsed_time(fe) / 5
    log(f"[bench] plain loop: {loop_rate:.1f} refactors/s (max eta {loop_eta:.2e}); "
        f"single factor {single_factor_ms:.3f} ms; batch/loop = {value / loop_rate:.1f}x")

    # ---------------- batched refactor alone -----------------------------
    fb = _events(lambda: plan.batch_factor(ax, lu=lu), 11)
    log(f"[bench] batched refactor (refresh + factor, no solve): {fb:.3f} ms -> {B / fb * 1e3:.1f} refactors/s")

    # ---------------- roofline of the batch step ------------------------
    per_mat = [alg_bytes_factor(an.nnz_lu, n, B) + int(it) * alg_bytes_refine(data.size, an.nnz_lu, n, B)
               for it in iters_h]
    traffic_alg = float(np.sum(per_mat))
    achieved = traffic_alg / (kernel_ms * 1e-3) / 1e9
    traffic = None
    prof = os.path.join(ROOT, "profiles", "batch_traffic.json")
    if os.path.exists(prof):
        with open(prof) as f:
            traffic = json.load(f).get("bytes_per_launch")

    # ---------------- side configurations --------------------------------
    small = uneven = grad = None
    if not args.no_extra:
        small = bench_small(dev, log)
        uneven = bench_uneven(dev, log, args.tol, args.maxit)
        grad = bench_autograd(dev, log)

    # ---------------- n = 500k single refactor --------------------------
    large = None if args.no_large else bench_large(dev, log, args.tol)

    # ---------------- CPU baseline (NICSLU, reference code) ------------
    cpu = None
    if not args.no_cpu:
        from oracle import ref
        if ref.available():
            cpu = cpu_baseline(ref, data, indices, indptr, n, ax_h, args, log)
        else:
            log("[bench] oracle/_ref not built: no CPU baseline")

    line = {
        "metric": "numeric refactor ms + batched refactors/s at <=1e-10 bwd err vs plain FP32",
        "value": round(value, 2),
        "unit": "refactors/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": round(wall / K * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"batched same-pattern refactor + FP64 refine: {B} MNA matrices, n={n}, "
                               f"nnz(A)={data.size}, nnz(L+U)={an.nnz_lu}, {an.num_lev} levels, tol {args.tol:g}",
                   "batch": B, "n": n, "nnz": int(data.size), "nnz_lu": an.nnz_lu, "levels": an.num_lev,
                   "parallelism": f"{plan.batch_grid_levels} wide steps as grid launches over 16-matrix "
                                  f"interleaved groups, then one workgroup per matrix"},
        "correct": ok,
        "refine_iters_max": int(iters_h.max()),
        "eta_max": float(eta_h.max()),
        "accuracy": batch_acc,
        "plain_fp32_loop": {"value": round(loop_rate, 2), "unit": "refactors/s", "sample": S,
                            "single_factor_ms": round(single_factor_ms, 4),
                            "speedup": round(value / loop_rate, 2)},
        "pcie_inclusive": {"value": round(pcie_rate, 2), "unit": "refactors/s"},
        "batch_refactor_only": {"value": round(B / fb * 1e3, 2), "unit": "refactors/s", "ms": round(fb, 4)},
        "small_factor_solve": small,
        "large_refactor": large,
        "uneven_group": uneven,
        "autograd": grad,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                     "kernel": "one batch step: k_il_zero_fill, k_il_scatter, k_factor_grid_il per wide "
                               "step, k_il_to_rows, k_batch_same_fused (per-matrix factor tail, sweep copy, refinement)",
                     "kernel_ms": round(kernel_ms, 4), "alg_bytes": round(traffic_alg)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
