"""GPU parity of the numeric phase against the oracle (C restatement of the
reference's FP32 arithmetic) — factorization, GESP, solves, refinement."""

import numpy as np
import pytest

from helpers import backward_error, entry_rel_err, mna_case

pytestmark = pytest.mark.gpu

CASES = ["mna600", "mna2k", "small", "add32", "batch"]


@pytest.fixture(scope="module")
def torch_mod():
    import torch
    return torch


def _setup(name):
    from oracle import oracle as O
    from pyglu.analysis import Analysis
    from pyglu.device import Plan
    data, indices, indptr, n = mna_case(name)
    an = Analysis(data, indices, indptr, n)
    return data, indices, indptr, n, an, Plan(an), O.SymPattern(an)


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("perturb", [False, True])
def test_factor_matches_oracle(torch_mod, name, perturb):
    """FP32 factors bit-identical to the oracle's sequential pass over each
    level's columns (a valid order of the reference's atomics), hence within
    the 1e-4-per-entry bar."""
    from oracle import oracle as O
    torch = torch_mod
    data, indices, indptr, n, an, plan, pat = _setup(name)
    print(f"[{name}] n={n} nnz_lu={an.nnz_lu} levels={an.num_lev} perturb={perturb}", flush=True)
    lu = torch.from_numpy(an.val.copy()).cuda()
    plan.factor(lu, perturb=perturb)
    torch.cuda.synchronize()
    got = lu.cpu().numpy()
    want, pert = O.factor_f32(pat, an.val, perturb=perturb)
    assert np.isfinite(got).all()
    err = entry_rel_err(got, want, pat.sym_c_ptr)
    exact = np.mean(got.view(np.uint32) == want.view(np.uint32))
    print(f"  max rel err {err.max():.3e}  exact fraction {exact:.6f}")
    assert err.max() <= 1e-4
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("name", ["mna2k", "small", "add32"])
def test_factor_host_matches_device(torch_mod, name):
    """glu_factor_host (the LUonDevice contract: host buffer in/out) and
    repeated glu_factor calls on device memory give identical bits."""
    torch = torch_mod
    data, indices, indptr, n, an, plan, pat = _setup(name)
    host = an.val.copy()
    plan.factor_host(host)
    for _ in range(3):
        lu = torch.from_numpy(an.val.copy()).cuda()
        plan.factor(lu)
        torch.cuda.synchronize()
        assert np.array_equal(host.view(np.uint32), lu.cpu().numpy().view(np.uint32))


@pytest.mark.parametrize("name", CASES + ["grad"])
def test_solve_bit_exact(torch_mod, name):
    """With the same factors, the device solve equals the reference's FP32
    host solve (restated in the oracle) bit for bit (n=50k "grad": ready
    bits in LDS, the vector in global memory)."""
    from oracle import oracle as O
    from pyglu import mna
    torch = torch_mod
    data, indices, indptr, n, an, plan, pat = _setup(name)
    lu = torch.from_numpy(an.val.copy()).cuda()
    plan.factor(lu)
    b = mna.rhs(n, 101).astype(np.float32)
    x = plan.solve(lu, torch.from_numpy(b).cuda())
    torch.cuda.synchronize()
    want = O.solve_f32(pat, lu.cpu().numpy(), an.col_perm, an.row_perm_inv, an.col_scale_perm, an.row_scale, b)
    assert np.array_equal(x.cpu().numpy().view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("name", ["mna2k", "small", "add32"])
def test_transposed_solve(torch_mod, name):
    """A' x = b with the FP32 factors: close to the FP64 oracle solve."""
    from oracle import oracle as O
    torch = torch_mod
    data, indices, indptr, n, an, plan, pat = _setup(name)
    lu = torch.from_numpy(an.val.copy()).cuda()
    plan.factor(lu)
    b = np.random.default_rng(3).uniform(-1, 1, n)
    x = plan.solve(lu, torch.from_numpy(b.astype(np.float32)).cuda(), transpose=True).cpu().numpy()
    lu64 = O.factor_f64(pat, O.lu_values_f64(pat, an.dump_ax, an.dump_ai, an.dump_ap))
    x64 = O.solve_f64(pat, lu64, an, b.astype(np.float32).astype(np.float64), transpose=True)
    rel = np.abs(x - x64).max() / np.abs(x64).max()
    print(f"[{name}] transposed solve rel err vs FP64 {rel:.2e}")
    assert rel < 1e-3


@pytest.mark.parametrize("name", ["mna2k", "small", "add32", "batch"])
@pytest.mark.parametrize("transpose", [False, True])
def test_refine_backward_error(torch_mod, name, transpose):
    """FP32 LU + FP64 residual refinement reaches eta <= 1e-10 (A and A')."""
    torch = torch_mod
    data, indices, indptr, n, an, plan, pat = _setup(name)
    lu = torch.from_numpy(an.val.copy()).cuda()
    plan.factor(lu)
    b = np.random.default_rng(5).uniform(-1, 1, n)
    x, it, eta = plan.refine(torch.from_numpy(data).cuda(), lu, torch.from_numpy(b).cuda(), tol=1e-10,
                             maxit=20, transpose=transpose)
    xh = x.cpu().numpy()
    if transpose:
        import scipy.sparse as sp
        At = sp.csc_matrix((data, indices, indptr), shape=(n, n)).T.tocsc()
        eta_h = backward_error(At.data, At.indices, At.indptr, n, xh, b)
    else:
        eta_h = backward_error(data, indices, indptr, n, xh, b)
    print(f"[{name}] transpose={transpose} iters={it} eta={eta:.2e} host eta={eta_h:.2e}")
    assert eta <= 1e-10 and eta_h <= 2e-10


def test_batch_matches_single(torch_mod):
    """The persistent same-pattern batch gives, per matrix, what the single
    path gives: refreshed values, factors within the atomics' freedom, and
    solutions refined to the tolerance."""
    from pyglu import mna
    torch = torch_mod
    data, indices, indptr, n, an, plan, pat = _setup("mna2k")
    B = 6
    ax = np.stack([mna.perturb_values(data, 3000 + b) for b in range(B)])
    bb = np.stack([mna.rhs(n, 3500 + b) for b in range(B)])
    axd = torch.from_numpy(ax).cuda()
    x, lu, iters, eta = plan.batch_solve(axd, torch.from_numpy(bb).cuda(), tol=1e-10, maxit=20)
    torch.cuda.synchronize()
    lu_ref = plan.refresh(axd)
    for b in range(B):
        one = lu_ref[b].clone()
        plan.factor(one)
        torch.cuda.synchronize()
        assert np.array_equal(lu[b].cpu().numpy().view(np.uint32), one.cpu().numpy().view(np.uint32))
        e = backward_error(ax[b], indices, indptr, n, x[b].cpu().numpy(), bb[b])
        print(f"  matrix {b}: iters={int(iters[b])} eta={float(eta[b]):.2e} host={e:.2e}")
        assert float(eta[b]) <= 1e-10 and e <= 2e-10


@pytest.mark.parametrize("name", ["small", "batch", "unsym"])
def test_refresh_matches_oracle(torch_mod, name):
    """glu_refresh and the batch kernels' value phase equal the oracle's
    restatement of NICSLU's refactor scaling (_I_NicsLU_MC64ScaleForRefact,
    (a * rs) * csp) + my_DumpA + predictLU bit for bit: on drifted value sets
    and on a value set where the analysis order a * (rs * csp) rounds to
    other FP32 values (the oracle with that order must then differ).  The
    batch kernels' phase 1 is checked through their factors, which must be
    the oracle's factorization of the oracle's refreshed values."""
    from helpers import order_sensitive_values, unsym_case
    from oracle import oracle as O
    from pyglu import mna
    from pyglu.analysis import Analysis
    from pyglu.device import Plan
    torch = torch_mod
    data, indices, indptr, n = unsym_case() if name == "unsym" else mna_case(name)
    an = Analysis(data, indices, indptr, n)
    plan, pat = Plan(an), O.SymPattern(an)
    adv, sensitive = order_sensitive_values(an, indices, indptr, data)
    sets = np.stack([mna.perturb_values(data, 3000 + b) for b in range(3)] + [adv])
    print(f"[{name}] n={n} order-sensitive entries {int(sensitive.sum())}", flush=True)
    assert sensitive.any()
    axd = torch.from_numpy(sets).cuda()
    got = plan.refresh(axd).cpu().numpy()
    lub = plan.batch_factor(axd).cpu().numpy()
    for b in range(sets.shape[0]):
        _, want = O.refresh(pat, an, indices, indptr, sets[b])
        assert np.array_equal(got[b].view(np.uint32), want.view(np.uint32)), f"refresh, set {b}"
        fw, _ = O.factor_f32(pat, want)
        assert np.array_equal(lub[b].view(np.uint32), fw.view(np.uint32)), f"batch factor, set {b}"
    _, wrong = O.refresh(pat, an, indices, indptr, adv, order=1)
    assert np.count_nonzero(got[-1].view(np.uint32) != wrong.view(np.uint32)) > 0


def test_group_matches_single(torch_mod):
    """An uneven group of distinct patterns, one fused launch."""
    from pyglu import mna
    from pyglu.analysis import Analysis
    from pyglu.device import Group, Plan
    torch = torch_mod
    mats, plans, axs, bs = [], [], [], []
    for spec, seed in mna.uneven_specs(6, seed0=4100):
        spec = mna.MNASpec(n_nodes=min(spec.n_nodes, 3000), n_vsrc=min(spec.n_vsrc, 20),
                           extra_hi=spec.extra_hi, n_rails=spec.n_rails, rail_deg=min(spec.rail_deg, 100))
        d, i, p, n = mna.mna_csc(spec, seed)
        an = Analysis(d, i, p, n)
        mats.append((d, i, p, n))
        plans.append(Plan(an))
        axs.append(torch.from_numpy(d).cuda())
        bs.append(torch.from_numpy(mna.rhs(n, seed + 1)).cuda())
    g = Group(plans)
    xs, lus, iters, eta = g.solve(axs, bs, tol=1e-10, maxit=20)
    torch.cuda.synchronize()
    for k, (d, i, p, n) in enumerate(mats):
        e = backward_error(d, i, p, n, xs[k].cpu().numpy(), bs[k].cpu().numpy())
        print(f"  group member {k}: n={n} iters={int(iters[k])} eta={float(eta[k]):.2e} host={e:.2e}")
        assert float(eta[k]) <= 1e-10 and e <= 2e-10


def test_group_split_same_bits(torch_mod):
    """The heaviest members of an uneven group moved to the single-matrix
    path beside the fused launch (glu_group_set_split 1 and 2) give the same
    factors, solutions, iteration counts and eta, bit for bit, as the fused
    launch alone (split 0), on two solves in a row."""
    from pyglu import mna
    from pyglu.analysis import Analysis
    from pyglu.device import Group, Plan
    torch = torch_mod
    plans, axs, bs = [], [], []
    for spec, seed in mna.uneven_specs(6, seed0=4200):
        spec = mna.MNASpec(n_nodes=min(spec.n_nodes, 4000), n_vsrc=min(spec.n_vsrc, 20),
                           extra_hi=spec.extra_hi, n_rails=spec.n_rails, rail_deg=min(spec.rail_deg, 120))
        d, i, p, n = mna.mna_csc(spec, seed)
        plans.append(Plan(Analysis(d, i, p, n)))
        axs.append(torch.from_numpy(mna.perturb_values(d, seed + 3)).cuda())
        bs.append(torch.from_numpy(mna.rhs(n, seed + 1)).cuda())
    runs = {}
    for split in (0, 1, 2):
        g = Group(plans, split=split)
        for rep in range(2):
            xs, lus, iters, eta = g.solve(axs, bs, tol=1e-10, maxit=20)
            torch.cuda.synchronize()
            runs[(split, rep)] = ([x.cpu().numpy() for x in xs], [lu.cpu().numpy() for lu in lus],
                                  iters.cpu().numpy(), eta.cpu().numpy())
        g.close()
    want = runs[(0, 0)]
    assert np.all(want[3] <= 1e-10)
    for key, got in runs.items():
        print(f"  split {key[0]} solve {key[1]}: iters {got[2].tolist()}", flush=True)
        for k in range(len(plans)):
            assert np.array_equal(got[0][k].view(np.uint64), want[0][k].view(np.uint64)), (key, k)
            assert np.array_equal(got[1][k].view(np.uint32), want[1][k].view(np.uint32)), (key, k)
        assert np.array_equal(got[2], want[2]) and np.array_equal(got[3].view(np.uint64), want[3].view(np.uint64)), key
    with pytest.raises(Exception):
        Group(plans, split=3)


def test_autograd_against_dense_fp64(torch_mod):
    """dL/dA (on A's pattern) and dL/db of L = |x - x*|^2 against the dense
    FP64 reference gradient and central finite differences."""
    import scipy.sparse as sp
    from pyglu import mna
    from pyglu.analysis import Analysis
    from pyglu.device import Plan, spsolve_autograd
    torch = torch_mod
    d, i, p, n = mna.mna_csc(mna.MNASpec(n_nodes=300, n_vsrc=6, n_rails=2, rail_deg=30), 51)
    an = Analysis(d, i, p, n)
    plan = Plan(an)
    rng = np.random.default_rng(5000)
    xstar = rng.standard_normal(n)
    b = rng.uniform(-1, 1, n)
    vals = torch.from_numpy(d.copy()).cuda().requires_grad_(True)
    bt = torch.from_numpy(b).cuda().requires_grad_(True)
    x = spsolve_autograd(vals, bt, plan)
    loss = ((x - torch.from_numpy(xstar).cuda()) ** 2).sum()
    loss.backward()
    gA = vals.grad.cpu().numpy()
    gb = bt.grad.cpu().numpy()
    # dense FP64 reference
    A = sp.csc_matrix((d, i, p), shape=(n, n)).toarray()
    xd = np.linalg.solve(A, b)
    lam = np.linalg.solve(A.T, 2 * (xd - xstar))
    cols = np.repeat(np.arange(n), np.diff(p))
    gA_ref = -lam[i] * xd[cols]
    relA = np.abs(gA - gA_ref).max() / np.abs(gA_ref).max()
    relb = np.abs(gb - lam).max() / np.abs(lam).max()
    print(f"grad rel err: A {relA:.2e}  b {relb:.2e}")
    assert relA <= 1e-6 and relb <= 1e-6
    # central finite differences on a few entries (FP64 dense solves)
    h = 1e-6
    for k in rng.choice(d.size, 10, replace=False):
        Ap, Am = A.copy(), A.copy()
        Ap[i[k], cols[k]] += h
        Am[i[k], cols[k]] -= h
        fp = ((np.linalg.solve(Ap, b) - xstar) ** 2).sum()
        fm = ((np.linalg.solve(Am, b) - xstar) ** 2).sum()
        fd = (fp - fm) / (2 * h)
        assert abs(fd - gA[k]) <= 1e-5 * max(abs(fd), 1e-3 * np.abs(gA).max())


def test_graph_replay_matches_eager(torch_mod):
    """Factor and solve replayed from captured hipGraphs: every replay, with
    new contents in the captured buffers and eager work between replays,
    bit-identical to the eager launches; new buffers trigger a recapture."""
    from pyglu import mna
    from pyglu.analysis import Analysis
    from pyglu.device import Plan
    torch = torch_mod
    d, i, p, n = mna_case("mna2k")
    an = Analysis(d, i, p, n)
    pg, pe = Plan(an, use_graph=True), Plan(an)
    lu_g = torch.empty(an.nnz_lu, dtype=torch.float32, device="cuda")
    b_g = torch.empty(n, dtype=torch.float32, device="cuda")
    x_g = torch.empty(n, dtype=torch.float32, device="cuda")
    for rep in range(4):
        ax = torch.from_numpy(mna.perturb_values(d, 900 + rep)).cuda()
        b = torch.from_numpy(mna.rhs(n, 950 + rep).astype(np.float32)).cuda()
        lu_e = pe.refresh(ax).reshape(-1)
        pe.factor(lu_e, perturb=rep % 2 == 1)
        x_e = pe.solve(lu_e, b)
        torch.cuda.synchronize()
        if rep == 3:   # other buffers: recapture
            lu_g = torch.empty_like(lu_g)
            x_g = torch.empty_like(x_g)
        pg.refresh(ax, lu_g)
        pg.factor(lu_g, perturb=rep % 2 == 1)
        b_g.copy_(b)
        pg.solve(lu_g, b_g, x_g)
        torch.cuda.synchronize()
        assert torch.equal(lu_g.view(torch.int32), lu_e.view(torch.int32)), rep
        assert torch.equal(x_g.view(torch.int32), x_e.view(torch.int32)), rep
        xt_e = pe.solve(lu_e, b, transpose=True)
        xt_g = pg.solve(lu_g, b_g, torch.empty_like(x_g), transpose=True)
        torch.cuda.synchronize()
        assert torch.equal(xt_g.view(torch.int32), xt_e.view(torch.int32)), rep


def test_batch_krylov_modes(torch_mod):
    """Every refinement mode of the batch (plain, FGMRES(1..4)) reaches the
    backward-error target on the same matrices; the factors do not depend on
    the mode, and the FGMRES results are bit-reproducible (fixed-order sums)."""
    from pyglu import mna
    from pyglu.device import Plan
    torch = torch_mod
    data, indices, indptr, n, an, plan, pat = _setup("mna2k")
    B = 5
    ax = np.stack([mna.perturb_values(data, 3100 + b) for b in range(B)])
    bb = np.stack([mna.rhs(n, 3600 + b) for b in range(B)])
    axd, bd = torch.from_numpy(ax).cuda(), torch.from_numpy(bb).cuda()
    lus = {}
    for kry in (-1, 1, 2, 3, 4):
        pk = Plan(an, krylov=kry)
        x, lu, iters, eta = pk.batch_solve(axd, bd, tol=1e-10, maxit=20)
        x2, _, iters2, _ = pk.batch_solve(axd, bd, tol=1e-10, maxit=20)
        torch.cuda.synchronize()
        print(f"  krylov={kry}: iters {iters.cpu().numpy().tolist()}", flush=True)
        assert np.all(eta.cpu().numpy() <= 1e-10)
        assert torch.equal(x.view(torch.int64), x2.view(torch.int64)) and torch.equal(iters, iters2)
        for b in range(B):
            assert backward_error(ax[b], indices, indptr, n, x[b].cpu().numpy(), bb[b]) <= 2e-10
        lus[kry] = lu.cpu().numpy().view(np.uint32)
    for kry in lus:
        assert np.array_equal(lus[kry], lus[-1])


def test_plan_rejects_bad_krylov(torch_mod):
    from pyglu.device import Plan
    data, indices, indptr, n, an, plan, pat = _setup("small")
    for bad in (-2, 5):
        with pytest.raises(RuntimeError):
            Plan(an, krylov=bad)
    with pytest.raises(RuntimeError):
        Plan(an, batch_grid_levels=-2)


@pytest.mark.parametrize("perturb", [False, True])
def test_batch_factor_matches_single(torch_mod, perturb):
    """glu_batch_factor (refresh + factor, one launch) gives, per matrix, the
    single path's factors bit for bit, and the same factors as the batched
    solve."""
    from pyglu import mna
    torch = torch_mod
    data, indices, indptr, n, an, plan, pat = _setup("mna2k")
    B = 7
    axd = torch.from_numpy(np.stack([mna.perturb_values(data, 3300 + b) for b in range(B)])).cuda()
    bd = torch.from_numpy(np.stack([mna.rhs(n, 3700 + b) for b in range(B)])).cuda()
    lu = plan.batch_factor(axd, perturb=perturb)
    _, lu_s, _, _ = plan.batch_solve(axd, bd, tol=1e-10, maxit=20, perturb=perturb)
    torch.cuda.synchronize()
    assert torch.equal(lu.view(torch.int32), lu_s.view(torch.int32))
    for b in range(B):
        one = plan.refresh(axd[b]).reshape(-1)
        plan.factor(one, perturb=perturb)
        torch.cuda.synchronize()
        assert torch.equal(one.view(torch.int32), lu[b].view(torch.int32)), b


@pytest.mark.parametrize("B", [40, 32, 17, 5, 1])
@pytest.mark.parametrize("perturb", [False, True])
def test_batch_grid_steps_same_bits(torch_mod, B, perturb):
    """The batched kernels' leading steps as grid launches over an
    interleaved copy of the batch (groups of 16 matrices; 40 = two full
    groups and a partial one, 32 = two full groups, whose copy is zeroed on
    the fill-in lines only, 17 = one full group and a single matrix) give
    the same bits as the workgroup-per-matrix
    kernel alone (one, three and the default number of grid steps):
    factors, solutions, iteration counts and backward errors of
    glu_batch_solve, factors of glu_batch_factor, and the single path's
    factors."""
    from pyglu import mna
    from pyglu.device import Plan
    torch = torch_mod
    data, indices, indptr, n, an, plan, pat = _setup("batch")
    axd = torch.from_numpy(np.stack([mna.perturb_values(data, 5100 + b) for b in range(B)])).cuda()
    bd = torch.from_numpy(np.stack([mna.rhs(n, 5300 + b) for b in range(B)])).cuda()
    assert plan.batch_grid_levels > 0
    runs = {}
    for opt in (-1, 0, 1, 3):
        pl = plan if opt == 0 else Plan(an, batch_grid_levels=opt)
        assert pl.batch_grid_levels == {-1: 0, 0: plan.batch_grid_levels, 1: 1, 3: 3}[opt]
        print(f"  B={B} perturb={perturb} batch_grid_levels={pl.batch_grid_levels}", flush=True)
        x, lu, iters, eta = pl.batch_solve(axd, bd, tol=1e-10, maxit=20, perturb=perturb)
        luf = pl.batch_factor(axd, perturb=perturb)
        torch.cuda.synchronize()
        runs[opt] = [t.clone() for t in (x, lu, iters, eta, luf)]
        if pl is not plan:
            pl.close()
    for opt in (0, 1, 3):
        for a, b in zip(runs[-1], runs[opt]):
            assert torch.equal(a.view(torch.uint8), b.view(torch.uint8)), opt
    assert torch.equal(runs[0][1].view(torch.int32), runs[0][4].view(torch.int32))
    one = plan.refresh(axd[B - 1]).reshape(-1)
    plan.factor(one, perturb=perturb)
    torch.cuda.synchronize()
    assert torch.equal(one.view(torch.int32), runs[0][1][B - 1].view(torch.int32))


@pytest.mark.parametrize("perturb", [False, True])
def test_batch_vector_beyond_lds(torch_mod, perturb):
    """A same-pattern batch whose order's correction vector does not fit in
    LDS (n = 50,000: 200 KB of FP32 against 160 KB per workgroup), so the
    refinement keeps its vector in global memory and only the ready bits in
    LDS (factor and refinement as two launches): the same bits with and
    without the grid steps, each matrix's factors equal to the single path's
    and matrix 0's to the oracle's factorization of the oracle's refreshed
    values, every backward error within 1e-10 (with GESP, whose replaced
    pivots can slow the refinement: within 1e-10 or the true residual's
    backward error reported at maxit)."""
    from oracle import oracle as O
    from pyglu import mna
    from pyglu.analysis import Analysis
    from pyglu.device import Plan
    torch = torch_mod
    data, indices, indptr, n = mna.mna_csc(mna.SPEC_GRAD, mna.SEED_GRAD)
    assert 4 * n > 160 * 1024
    an = Analysis(data, indices, indptr, n)
    pat = O.SymPattern(an)
    plan, none = Plan(an), Plan(an, batch_grid_levels=-1)
    assert plan.batch_grid_levels > 1
    B = 3
    ax = np.stack([mna.perturb_values(data, 6100 + b) for b in range(B)])
    bb = np.stack([mna.rhs(n, 6300 + b) for b in range(B)])
    axd, bd = torch.from_numpy(ax).cuda(), torch.from_numpy(bb).cuda()
    runs = []
    for pl in (none, plan):
        x, lu, iters, eta = pl.batch_solve(axd, bd, tol=1e-10, maxit=20, perturb=perturb)
        luf = pl.batch_factor(axd, perturb=perturb)
        torch.cuda.synchronize()
        runs.append([t.clone() for t in (x, lu, iters, eta, luf)])
    none.close()
    for a, b in zip(runs[0], runs[1]):
        assert torch.equal(a.view(torch.uint8), b.view(torch.uint8))
    x, lu, iters, eta, luf = runs[1]
    assert torch.equal(lu.view(torch.int32), luf.view(torch.int32))
    for b in range(B):
        one = plan.refresh(axd[b]).reshape(-1)
        plan.factor(one, perturb=perturb)
        torch.cuda.synchronize()
        assert torch.equal(one.view(torch.int32), lu[b].view(torch.int32)), b
        e = backward_error(ax[b], indices, indptr, n, x[b].cpu().numpy(), bb[b])
        print(f"  n={n} matrix {b}: iters={int(iters[b])} eta={float(eta[b]):.2e} host={e:.2e}", flush=True)
        if perturb:
            # replaced pivots make the factors a weaker preconditioner: the
            # refinement may stop at maxit, reporting the true residual's eta
            assert abs(float(eta[b]) - e) <= 0.01 * e and (float(eta[b]) <= 1e-10 or int(iters[b]) == 20)
        else:
            assert float(eta[b]) <= 1e-10 and e <= 2e-10
    _, ref = O.refresh(pat, an, indices, indptr, ax[0])
    want, _ = O.factor_f32(pat, ref, perturb=perturb)
    assert np.array_equal(lu[0].cpu().numpy().view(np.uint32), want.view(np.uint32))
    plan.close()


def test_batch_grid_steps_replaced_pivots(torch_mod):
    """GESP pivot replacement inside the grid steps: value sets whose
    columns are in part scaled by 1e-12 (the frozen analysis scales leave
    those pivots below the GESP threshold, early columns among them) give
    the same bits with and without the grid steps, the oracle's perturbed
    factorization bit for bit, and not the oracle's unperturbed one."""
    from oracle import oracle as O
    from pyglu import mna
    from pyglu.device import Plan
    torch = torch_mod
    data, indices, indptr, n, an, plan, pat = _setup("batch")
    rng = np.random.default_rng(77)
    B = 18
    sets = []
    for b in range(B):
        v = mna.perturb_values(data, 5600 + b).copy()
        for c in rng.choice(n, 600, replace=False):
            v[indptr[c]:indptr[c + 1]] *= 1e-12
        sets.append(v)
    axd = torch.from_numpy(np.stack(sets)).cuda()
    bd = torch.from_numpy(np.stack([mna.rhs(n, 5700 + b) for b in range(B)])).cuda()
    assert plan.batch_grid_levels > 1
    none = Plan(an, batch_grid_levels=-1)
    runs = []
    for pl in (none, plan):
        x, lu, iters, eta = pl.batch_solve(axd, bd, tol=1e-10, maxit=5, perturb=True)
        luf = pl.batch_factor(axd, perturb=True)
        torch.cuda.synchronize()
        runs.append([t.clone() for t in (x, lu, iters, eta, luf)])
    none.close()
    for a, b in zip(runs[0], runs[1]):
        assert torch.equal(a.view(torch.uint8), b.view(torch.uint8))
    _, ref = O.refresh(pat, an, indices, indptr, sets[0])
    want, _ = O.factor_f32(pat, ref, perturb=True)
    plain, _ = O.factor_f32(pat, ref, perturb=False)
    got = runs[1][4][0].cpu().numpy()
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    assert not np.array_equal(plain.view(np.uint32), want.view(np.uint32))


def test_abft_device_matches_oracle(torch_mod):
    """glu_abft_sums / glu_abft_check (src/symbolic.cc:163-197) on a batch of
    three value sets: checksums bit-identical to the oracle's, the same
    verdict and largest difference on correct factors, and the planted
    error's column reported in the corrupted member only."""
    from oracle import oracle as O
    from pyglu import mna
    torch = torch_mod
    data, indices, indptr, n, an, plan, pat = _setup("mna2k")
    B = 3
    axd = torch.from_numpy(np.stack([mna.perturb_values(data, 3900 + b) for b in range(B)])).cuda()
    val0 = torch.stack([plan.refresh(axd[b]).reshape(-1) for b in range(B)])
    lu = val0.clone()
    for b in range(B):
        plan.factor(lu[b])
    cca = plan.abft_sums(val0)
    c = n // 2
    lu[1, int(pat.l_col_ptr[c])] += 1.0
    first, mx = plan.abft_check(lu, cca)
    torch.cuda.synchronize()
    v0h, luh, ccah = val0.cpu().numpy(), lu.cpu().numpy(), cca.cpu().numpy()
    fh, mh = first.cpu().numpy(), mx.cpu().numpy()
    for b in range(B):
        want = O.abft_cca(pat, v0h[b])
        assert np.array_equal(ccah[b].view(np.uint32), want.view(np.uint32))
        f, _, m = O.abft_check(pat, luh[b], want)
        assert int(fh[b]) == f and mh[b] == np.float32(m), (b, int(fh[b]), f, mh[b], m)
    assert int(fh[0]) == -1 and int(fh[2]) == -1 and int(fh[1]) == c


def test_autotune(torch_mod):
    """The autotuner times every candidate, returns the fastest, and a plan
    built from its choice gives the default plan's factors bit for bit (the
    schedule parameters never change results)."""
    from pyglu import mna
    from pyglu.autotune import report, tune_batch, tune_factor
    from pyglu.device import Plan
    torch = torch_mod
    data, indices, indptr, n, an, plan, pat = _setup("mna2k")
    res = tune_factor(an, fuse=(256, 2048), threads=(512, 1024), windows=(0, -1), reps=2)
    print(report(res), flush=True)
    assert len(res.table) == 5 and res.ms == min(r[3] for r in res.table)
    lu0 = torch.from_numpy(an.val.copy()).cuda()
    plan.factor(lu0)
    tuned = res.plan(an)
    lu1 = torch.from_numpy(an.val.copy()).cuda()
    tuned.factor(lu1)
    torch.cuda.synchronize()
    assert torch.equal(lu0.view(torch.int32), lu1.view(torch.int32))
    B = 4
    axd = torch.from_numpy(np.stack([mna.perturb_values(data, 3200 + b) for b in range(B)])).cuda()
    bd = torch.from_numpy(np.stack([mna.rhs(n, 3600 + b) for b in range(B)])).cuda()
    rb = tune_batch(an, axd, bd, threads=(512, 1024), krylov=(-1, 1), reps=1)
    print(report(rb), flush=True)
    assert len(rb.table) == 4 and rb.ms == min(r[3] for r in rb.table)
    x, lu, iters, eta = rb.plan(an).batch_solve(axd, bd, tol=1e-10, maxit=20)
    torch.cuda.synchronize()
    assert np.all(eta.cpu().numpy() <= 1e-10)
