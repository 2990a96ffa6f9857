"""Parity of the LDS-window factorization schedule and of the largest
configuration against the oracle.

* Plans built with a small LDS budget per window (lds_window_values), so
  that every test pattern gets at least three windows: single-matrix and
  batched factors bit-exact with the oracle, with and without GESP.
* The n=500k MNA matrix with the default plan (about ten windows): factors
  bit-exact with the oracle (+-GESP), FP64 refinement to eta <= 1e-10, and
  its forward error against the FP64 oracle solve.

The reference's schedule these replace is the level loop of
src/numeric.cu:551-617 (every level a launch, values in global memory)."""

import numpy as np
import pytest

from helpers import backward_error, entry_rel_err, mna_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_mod():
    import torch
    return torch


def _windowed_plan(an, want=3):
    """A plan whose factorization runs in at least `want` LDS windows: the
    largest per-window budget (values) that gives that many."""
    from pyglu.device import Plan
    for budget in (8192, 4096, 2048, 1024, 512, 256):
        plan = Plan(an, lds_window_values=budget)
        if plan.lds_window_count >= want:
            return plan, budget
        plan.close()
    raise AssertionError("no LDS budget gave enough windows")


@pytest.mark.parametrize("name", ["mna2k", "small", "batch"])
@pytest.mark.parametrize("perturb", [False, True])
def test_factor_many_windows_matches_oracle(torch_mod, name, perturb):
    from oracle import oracle as O
    from pyglu import mna
    from pyglu.analysis import Analysis
    torch = torch_mod
    data, indices, indptr, n = mna_case(name)
    an = Analysis(data, indices, indptr, n)
    pat = O.SymPattern(an)
    plan, budget = _windowed_plan(an)
    print(f"[{name}] n={n} levels={an.num_lev} budget={budget} windows={plan.lds_window_count} "
          f"steps in windows={plan.lds_window_steps} largest={plan.lds_window_max_values} perturb={perturb}",
          flush=True)
    lu = torch.from_numpy(an.val.copy()).cuda()
    plan.factor(lu, perturb=perturb)
    torch.cuda.synchronize()
    want, _ = O.factor_f32(pat, an.val, perturb=perturb)
    got = lu.cpu().numpy()
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), \
        f"single path: max rel err {entry_rel_err(got, want, pat.sym_c_ptr).max():.2e}"
    # the batched workgroup factorization runs the same windows
    sets = np.stack([mna.perturb_values(data, 3000 + b) for b in range(3)])
    axd = torch.from_numpy(sets).cuda()
    lub = plan.batch_factor(axd, perturb=perturb).cpu().numpy()
    for b in range(sets.shape[0]):
        _, v = O.refresh(pat, an, indices, indptr, sets[b])
        fw, _ = O.factor_f32(pat, v, perturb=perturb)
        assert np.array_equal(lub[b].view(np.uint32), fw.view(np.uint32)), f"batch factor {b}"
    if not perturb:
        bb = np.stack([mna.rhs(n, 3500 + b) for b in range(3)])
        x, lus, iters, eta = plan.batch_solve(axd, torch.from_numpy(bb).cuda(), tol=1e-10, maxit=20)
        assert torch.equal(lus.view(torch.int32), torch.from_numpy(lub).cuda().view(torch.int32))
        for b in range(3):
            e = backward_error(sets[b], indices, indptr, n, x[b].cpu().numpy(), bb[b])
            assert float(eta[b]) <= 1e-10 and e <= 2e-10, (b, float(eta[b]), e)
    plan.close()


def test_large_mna_default_plan(torch_mod):
    """n=500,000 MNA (3.4M nnz, 15.4M nnz(L+U), 3,606 levels), default plan."""
    import time

    from oracle import oracle as O
    from pyglu import mna
    from pyglu.analysis import Analysis
    from pyglu.device import Plan
    torch = torch_mod
    t0 = time.perf_counter()
    d, i, p, n = mna.mna_csc(mna.SPEC_LARGE, mna.SEED_LARGE)
    an = Analysis(d, i, p, n)
    plan = Plan(an)
    pat = O.SymPattern(an)
    print(f"n={n} nnz={d.size} nnz_lu={an.nnz_lu} levels={an.num_lev} windows={plan.lds_window_count} "
          f"(steps {plan.lds_window_steps}, largest {plan.lds_window_max_values} values); "
          f"analysis + plan {time.perf_counter() - t0:.1f} s", flush=True)
    assert plan.lds_window_count >= 2
    for perturb in (False, True):
        lu = torch.from_numpy(an.val.copy()).cuda()
        plan.factor(lu, perturb=perturb)
        torch.cuda.synchronize()
        want, _ = O.factor_f32_sparse(pat, an.val, perturb=perturb)
        got = lu.cpu().numpy()
        err = entry_rel_err(got, want, pat.sym_c_ptr).max()
        exact = np.array_equal(got.view(np.uint32), want.view(np.uint32))
        print(f"  perturb={perturb}: bit-exact={exact} max rel err {err:.2e}", flush=True)
        assert exact
    lu = torch.from_numpy(an.val.copy()).cuda()
    plan.factor(lu)
    b = mna.rhs(n, 102)
    ax = torch.from_numpy(d).cuda()
    x, it, eta = plan.refine(ax, lu, torch.from_numpy(b).cuda(), tol=1e-10, maxit=20)
    xh = x.cpu().numpy()
    eta_h = backward_error(d, i, p, n, xh, b)
    lu64 = O.factor_f64(pat, O.lu_values_f64(pat, an.dump_ax, an.dump_ai, an.dump_ap))
    x64 = O.solve_f64(pat, lu64, an, b)
    fwd = float(np.abs(xh - x64).max() / np.abs(x64).max())
    x32 = plan.solve(lu, torch.from_numpy(b.astype(np.float32)).cuda()).cpu().numpy().astype(np.float64)
    fwd32 = float(np.abs(x32 - x64).max() / np.abs(x64).max())
    print(f"  refine: {it} solves, eta {eta:.2e} (host {eta_h:.2e}); forward error vs FP64 oracle: "
          f"refined {fwd:.2e}, FP32 solve {fwd32:.2e}", flush=True)
    assert eta <= 1e-10 and eta_h <= 2e-10
    assert np.isfinite(fwd) and fwd < 1e-3
    plan.close()
