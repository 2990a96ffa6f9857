"""Full-size GPU checks: the uneven group (64 distinct MNA patterns, n in
[1k, 50k]), the differentiable solve at n=50,000 and the same-pattern batch
of 256 matrices at n=20,000."""

import numpy as np
import pytest

from helpers import backward_error

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_mod():
    import torch
    return torch


def test_uneven_group_full(torch_mod):
    """One fused launch over 64 patterns: every member refined to the
    backward-error target; the factors of 8 members (the smallest, the
    largest, n=37,815 with its several LDS windows among them) bit-identical
    to the oracle's factorization of the oracle's refactor values, and of
    three members to the single-matrix path."""
    from oracle import oracle as O
    from pyglu import mna
    from pyglu.analysis import Analysis
    from pyglu.device import Group, Plan
    torch = torch_mod
    mats, plans, axs, bs, ans = [], [], [], [], []
    for m, (spec, seed) in enumerate(mna.uneven_specs()):
        d, i, p, n = mna.mna_csc(spec, seed)
        mats.append((d, i, p, n))
        ans.append(Analysis(d, i, p, n))
        plans.append(Plan(ans[-1]))
        axs.append(torch.from_numpy(mna.perturb_values(d, seed + 7)).cuda())
        bs.append(torch.from_numpy(mna.rhs(n, 4500 + m)).cuda())
    print(f"64 patterns: n {min(p.n for p in plans)}..{max(p.n for p in plans)}, "
          f"levels up to {max(p.num_lev for p in plans)}", flush=True)
    g = Group(plans)
    xs, lus, iters, eta = g.solve(axs, bs, tol=1e-10, maxit=20)
    torch.cuda.synchronize()
    eta_h = eta.cpu().numpy()
    assert np.all(eta_h <= 1e-10), eta_h.max()
    for k, (d, i, p, n) in enumerate(mats):
        e = backward_error(axs[k].cpu().numpy(), i, p, n, xs[k].cpu().numpy(), bs[k].cpu().numpy())
        assert e <= 2e-10, (k, n, e)
    order = np.argsort([p.n for p in plans])
    for k in (int(order[0]), int(order[len(order) // 2]), int(order[-1])):
        lu = plans[k].refresh(axs[k]).reshape(-1)
        plans[k].factor(lu)
        assert torch.equal(lu.view(torch.int32), lus[k].view(torch.int32)), k
    picks = [int(order[q]) for q in np.linspace(0, len(order) - 1, 8).round().astype(int)]
    for k in picks:
        d, i, p, n = mats[k]
        pat = O.SymPattern(ans[k])
        _, v = O.refresh(pat, ans[k], i, p, axs[k].cpu().numpy())
        want, _ = O.factor_f32_sparse(pat, v)
        got = lus[k].cpu().numpy()
        print(f"  member {k}: n={n} windows={plans[k].lds_window_count} oracle bit-exact "
              f"{np.array_equal(got.view(np.uint32), want.view(np.uint32))}", flush=True)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (k, n)
    g.close()


def test_autograd_full_size(torch_mod):
    """loss = |x - x*|^2 through spsolve_autograd at n=50,000: gradients for
    A's stored values and b within 1e-6 (normwise relative) of the sparse
    FP64 oracle's adjoint gradient, and central finite differences (h=1e-6)
    of the FP64 loss on 100 random entries, each perturbed solve taken
    exactly by the Sherman-Morrison formula over the FP64 factors."""
    from oracle import oracle as O
    from pyglu import mna
    from pyglu.analysis import Analysis
    from pyglu.device import Plan, spsolve_autograd
    torch = torch_mod
    d, i, p, n = mna.mna_csc(mna.SPEC_GRAD, mna.SEED_GRAD)
    an = Analysis(d, i, p, n)
    plan = Plan(an)
    xstar = np.random.default_rng(5000).standard_normal(n)
    b = mna.rhs(n, 5001)
    vals = torch.from_numpy(d.copy()).cuda().requires_grad_(True)
    bt = torch.from_numpy(b.copy()).cuda().requires_grad_(True)
    x = spsolve_autograd(vals, bt, plan)
    loss = ((x - torch.from_numpy(xstar).cuda()) ** 2).sum()
    loss.backward()
    gA = vals.grad.cpu().numpy()
    gb = bt.grad.cpu().numpy()

    pat = O.SymPattern(an)
    lu64 = O.factor_f64(pat, O.lu_values_f64(pat, an.dump_ax, an.dump_ai, an.dump_ap))
    x64 = O.solve_f64(pat, lu64, an, b)
    lam = O.solve_f64(pat, lu64, an, 2 * (x64 - xstar), transpose=True)
    cols = np.repeat(np.arange(n), np.diff(p))
    gA_ref = -lam[i] * x64[cols]
    relA = np.abs(gA - gA_ref).max() / np.abs(gA_ref).max()
    relb = np.abs(gb - lam).max() / np.abs(lam).max()
    print(f"n={n}: grad rel err A {relA:.2e}  b {relb:.2e}", flush=True)
    assert relA <= 1e-6 and relb <= 1e-6

    h = 1e-6
    gmax = np.abs(gA).max()
    for k in np.random.default_rng(5002).choice(d.size, 100, replace=False):
        r, c = int(i[k]), int(cols[k])
        e_r = np.zeros(n)
        e_r[r] = 1.0
        z = O.solve_f64(pat, lu64, an, e_r)          # A^-1 e_r
        f = []
        for s in (h, -h):                             # (A + s e_r e_c^T)^-1 b
            xp = x64 - (s * x64[c] / (1.0 + s * z[c])) * z
            f.append(((xp - xstar) ** 2).sum())
        fd = (f[0] - f[1]) / (2 * h)
        assert abs(fd - gA[k]) <= 1e-5 * max(abs(fd), 1e-3 * gmax), (k, fd, gA[k])


def test_batch_full_size(torch_mod):
    """256 same-pattern n=20,000 MNA matrices in one launch: every matrix
    refined to the backward-error target (device eta and a host recheck),
    and the factors of three drifted matrices bit-identical to the
    single-matrix path and to the oracle's factorization of the oracle's
    refactor values."""
    from pyglu import mna
    from pyglu.analysis import Analysis
    from pyglu.device import Plan
    torch = torch_mod
    d, i, p, n = mna.mna_csc(mna.SPEC_BATCH, mna.SEED_BATCH)
    an = Analysis(d, i, p, n)
    plan = Plan(an)
    B = 256
    ax = np.stack([mna.perturb_values(d, 3000 + b) for b in range(B)])
    bb = np.stack([mna.rhs(n, 3500 + b) for b in range(B)])
    axd = torch.from_numpy(ax).cuda()
    x, lu, iters, eta = plan.batch_solve(axd, torch.from_numpy(bb).cuda(), tol=1e-10, maxit=20)
    torch.cuda.synchronize()
    eta_h = eta.cpu().numpy()
    it_h = iters.cpu().numpy()
    print(f"B={B}: iterations {np.bincount(it_h)}, eta max {eta_h.max():.2e}", flush=True)
    assert np.all(eta_h <= 1e-10), (int(np.argmax(eta_h)), eta_h.max())
    xh = x.cpu().numpy()
    for b in range(B):
        e = backward_error(ax[b], i, p, n, xh[b], bb[b])
        assert e <= 2e-10, (b, e)
    for b in (0, 97, 255):
        one = plan.refresh(axd[b]).reshape(-1)
        plan.factor(one)
        assert torch.equal(one.view(torch.int32), lu[b].view(torch.int32)), b
    # the factors of drifted matrices against the oracle directly
    from oracle import oracle as O
    pat = O.SymPattern(an)
    for b in (0, 97, 255):
        _, v = O.refresh(pat, an, i, p, ax[b])
        want, _ = O.factor_f32(pat, v)
        assert np.array_equal(lu[b].cpu().numpy().view(np.uint32), want.view(np.uint32)), b


def _fp64_solution(an, pat, indices, indptr, values, b):
    """x* from the FP64 oracle: FP64 refactor values (NICSLU's refactor
    scaling, dumped), FP64 LU on the same pattern, FP64 solve."""
    from oracle import oracle as O
    dump, _ = O.refresh(pat, an, indices, indptr, values)
    lu64 = O.factor_f64(pat, O.lu_values_f64(pat, dump, an.dump_ai, an.dump_ap))
    return O.solve_f64(pat, lu64, an, b)


def test_refinement_iterates_to_tight_tolerance(torch_mod):
    """With tol = 1e-13 the refinement must iterate (the FP32 solve alone
    stops near eta ~ 1e-11 on these matrices), and the refined forward error
    against the FP64 oracle must fall far below the FP32 solve's: the n=4,960
    matrix through glu_refine, and 16 matrices of the n=20,000 batch through
    the batched FGMRES refinement."""
    from oracle import oracle as O
    from pyglu import mna
    from pyglu.analysis import Analysis
    from pyglu.device import Plan
    torch = torch_mod
    tol = 1e-13
    d, i, p, n = mna.mna_csc(mna.SPEC_SMALL, mna.SEED_SMALL)
    an = Analysis(d, i, p, n)
    plan, pat = Plan(an), O.SymPattern(an)
    b = mna.rhs(n, 101)
    lu = plan.refresh(torch.from_numpy(d).cuda()).reshape(-1)
    plan.factor(lu)
    x32 = plan.solve(lu, torch.from_numpy(b.astype(np.float32)).cuda()).cpu().numpy().astype(np.float64)
    x, it, eta = plan.refine(torch.from_numpy(d).cuda(), lu, torch.from_numpy(b).cuda(), tol=tol, maxit=30)
    xs = _fp64_solution(an, pat, i, p, d, b)
    f32e = np.abs(x32 - xs).max() / np.abs(xs).max()
    fre = np.abs(x.cpu().numpy() - xs).max() / np.abs(xs).max()
    eh = backward_error(d, i, p, n, x.cpu().numpy(), b)
    print(f"n={n}: {it} solves, eta {eta:.2e} (host {eh:.2e}); forward error FP32 {f32e:.2e} refined {fre:.2e}",
          flush=True)
    assert it > 1 and eta <= tol and eh <= 2 * tol
    assert fre <= 1e-2 * f32e

    d, i, p, n = mna.mna_csc(mna.SPEC_BATCH, mna.SEED_BATCH)
    an = Analysis(d, i, p, n)
    plan, pat = Plan(an), O.SymPattern(an)
    B = 16
    ax = np.stack([mna.perturb_values(d, 3000 + k) for k in range(B)])
    bb = np.stack([mna.rhs(n, 3500 + k) for k in range(B)])
    axd = torch.from_numpy(ax).cuda()
    x, lus, iters, etas = plan.batch_solve(axd, torch.from_numpy(bb).cuda(), tol=tol, maxit=30)
    torch.cuda.synchronize()
    it_h, eta_h, xh = iters.cpu().numpy(), etas.cpu().numpy(), x.cpu().numpy()
    for k in range(B):
        one = lus[k].clone()
        x32 = plan.solve(one, torch.from_numpy(bb[k].astype(np.float32)).cuda()).cpu().numpy().astype(np.float64)
        xs = _fp64_solution(an, pat, i, p, ax[k], bb[k])
        f32e = np.abs(x32 - xs).max() / np.abs(xs).max()
        fre = np.abs(xh[k] - xs).max() / np.abs(xs).max()
        eh = backward_error(ax[k], i, p, n, xh[k], bb[k])
        print(f"  batch {k}: {it_h[k]} solves, eta {eta_h[k]:.2e} (host {eh:.2e}); forward error FP32 {f32e:.2e} "
              f"refined {fre:.2e}", flush=True)
        assert it_h[k] > 1 and eta_h[k] <= tol and eh <= 2 * tol
        assert fre <= 1e-2 * f32e
