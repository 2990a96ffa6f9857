"""glu_refine_noalloc (the refinement with its scratch in a static device
array, for callers that cannot allocate) against glu_refine, through the
C-ABI: the same solution bits, iteration count and eta on two calls in a
row, each on a smaller matrix than the one before (the array keeps the
previous call's contents), for A x = b and A' x = b, through both of
glu_refine's paths; and the size limit rejected with a clear error."""

import ctypes as C

import numpy as np
import pytest

from helpers import mna_case

pytestmark = pytest.mark.gpu


def _refine(fn, plan, ax, lu, b, transpose, tol=1e-12, maxit=20):
    import torch
    from pyglu import _capi as capi
    x = torch.full_like(b, float("nan"))
    it = C.c_int(-1)
    eta = C.c_double(-1.0)
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = getattr(capi.lib(), fn)(plan.handle, C.c_void_p(ax.data_ptr()), C.c_void_p(lu.data_ptr()),
                                 C.c_void_p(b.data_ptr()), C.c_void_p(x.data_ptr()), float(tol), int(maxit),
                                 int(bool(transpose)), C.byref(it), C.byref(eta), st)
    capi.check(rc, fn)
    torch.cuda.synchronize()
    return x.cpu().numpy(), it.value, eta.value


@pytest.mark.parametrize("transpose", [False, True])
def test_noalloc_matches_refine(transpose):
    import torch
    from pyglu import mna
    from pyglu.analysis import Analysis
    from pyglu.device import Plan
    # n = 50,000 (sweep vector beyond LDS: dataflow sweeps in one workgroup,
    # FP64 work across the grid), then the smaller n = 4,960 and n = 2,000
    # (everything in one workgroup)
    for name in ("grad", "small", "mna2k"):
        d, i, p, n = mna_case(name)
        plan = Plan(Analysis(d, i, p, n))
        ax = torch.from_numpy(mna.perturb_values(d, 11)).cuda()
        lu = plan.refresh(ax).reshape(-1)
        plan.factor(lu)
        b = torch.from_numpy(mna.rhs(n, 12)).cuda()
        print(f"[{name}] n={n} transpose={transpose}", flush=True)
        got = _refine("glu_refine_noalloc", plan, ax, lu, b, transpose)
        want = _refine("glu_refine", plan, ax, lu, b, transpose)
        print(f"  noalloc: {got[1]} solves eta {got[2]:.3e}; glu_refine: {want[1]} solves eta {want[2]:.3e}",
              flush=True)
        assert got[1] == want[1] and got[2] == want[2] and got[2] <= 1e-12
        assert np.array_equal(got[0].view(np.uint64), want[0].view(np.uint64))
        plan.close()


def test_noalloc_rejects_beyond_limit():
    """An order whose refinement scratch exceeds the static array is
    refused with GLU_ERR_ARGUMENT and a message naming the sizes."""
    import scipy.sparse as sp
    import torch
    from pyglu import _capi as capi
    from pyglu.analysis import Analysis
    from pyglu.device import Plan
    # n = 1M in dense 16 x 16 diagonal blocks (16M stored entries, no fill):
    # the sweeps' ready bits fit LDS, the refinement's scratch (the sweep
    # copy of 16M factor values alone is 64 MB) does not fit 64 MiB
    n, k = 1_000_000, 16
    rng = np.random.default_rng(1)
    indptr = np.arange(0, n * k + 1, k, dtype=np.int64)
    indices = ((np.arange(n) // k * k)[:, None] + np.arange(k)[None, :]).reshape(-1)
    data = rng.uniform(-1.0, 1.0, n * k)
    data[indices == np.repeat(np.arange(n), k)] += 2.0 * k   # diagonally dominant
    A = sp.csc_matrix((data, indices, indptr), shape=(n, n))
    plan = Plan(Analysis(A.data, A.indices.astype(np.int32), A.indptr.astype(np.int32), n))
    ax = torch.from_numpy(A.data.copy()).cuda()
    lu = plan.refresh(ax).reshape(-1)
    plan.factor(lu)
    b = torch.ones(n, dtype=torch.float64, device="cuda")
    with pytest.raises(capi.GluError) as e:
        _refine("glu_refine_noalloc", plan, ax, lu, b, False)
    assert capi.STATUS_NAMES[e.value.status] == "GLU_ERR_ARGUMENT" and "scratch" in str(e.value)
    x, it, eta = _refine("glu_refine", plan, ax, lu, b, False)   # the allocating entry point still serves it
    assert eta <= 1e-12
    plan.close()
