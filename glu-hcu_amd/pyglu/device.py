"""Device-level API over the C-ABI: plans, factor/solve/refine, batches,
uneven groups and the autograd wrapper of spsolve.

Device memory and streams come from PyTorch (torch.cuda tensors on the
current stream); every computation runs in libglu_hip.so.  There is no CPU
path: without a HIP device these calls raise.
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _capi as capi
from .analysis import Analysis


def _require_device():
    if not torch.cuda.is_available():
        raise RuntimeError("pyglu needs a HIP device (torch.cuda.is_available() is False)")


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _dptr(t: torch.Tensor):
    return C.c_void_p(t.data_ptr())


def _check_dev(t: torch.Tensor, dtype, numel, name):
    if not isinstance(t, torch.Tensor) or t.device.type != "cuda":
        raise TypeError(f"{name} must be a CUDA tensor")
    if t.dtype != dtype:
        raise TypeError(f"{name} must be {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if t.numel() != numel:
        raise ValueError(f"{name} has {t.numel()} elements, expected {numel}")


class Plan:
    """A device plan (``glu_plan``) for one sparsity pattern.

    ``source`` is an :class:`Analysis` (full plan: factor, solve, refresh,
    refine, batch) or any object exposing the Symbolic_Matrix arrays
    (factor-only plan, the LUonDevice contract).  ``use_graph`` replays the
    factor and solve launch sequences from captured hipGraphs (captured
    again when called with other buffers).  ``lds_windows``: how many runs of
    narrow levels, counted from the last, are factored with their values
    staged in LDS (0 = default, 64; -1 = none).  ``lds_window_values``: the
    most values one window may stage (0 = what the workgroup's LDS holds;
    smaller values give more, shorter windows).  ``solve_dataflow``: narrow
    levels of the single-matrix solves barrier-free (0 = default, 1 = on,
    -1 = off).  ``krylov``: refinement of the batched and grouped solves
    after the first solve (0 = default, restarted FGMRES(2); -1 = plain
    refinement; 1..4 = FGMRES(m)).  ``batch_grid_levels``: leading
    factorization steps of the batched kernels run as one grid launch per
    step over the whole batch (0 = default, the wide steps ahead of the
    first LDS window; -1 = none; k > 0 = at most k); the same bits either way.

    After creation, ``lds_window_count``, ``lds_window_steps`` and
    ``lds_window_max_values`` describe the factorization's LDS windows and
    ``batch_grid_levels`` the batched kernels' leading grid steps.
    """

    def __init__(self, source, fuse_units: int = 0, block_threads: int = 0, use_graph: bool = False,
                 lds_windows: int = 0, solve_dataflow: int = 0, krylov: int = 0, lds_window_values: int = 0,
                 batch_grid_levels: int = 0):
        _require_device()
        L = capi.lib()
        opts = capi.PlanOpts(fuse_units, block_threads, int(bool(use_graph)), int(lds_windows),
                             int(solve_dataflow), int(krylov), int(lds_window_values), int(batch_grid_levels))
        h = C.c_void_p()
        if isinstance(source, Analysis):
            capi.check(L.glu_plan_create_from_analysis(source.handle, C.byref(opts), C.byref(h)),
                       "glu_plan_create_from_analysis")
            self.analysis = source
        else:
            keep = [np.ascontiguousarray(getattr(source, k), dtype=np.uint32) for k in
                    ("sym_c_ptr", "sym_r_idx", "l_col_ptr", "csr_r_ptr", "csr_c_idx", "csr_diag_ptr")]
            keep += [np.ascontiguousarray(source.level_ptr, dtype=np.int32),
                     np.ascontiguousarray(source.level_idx, dtype=np.int32)]
            pat = capi.Pattern(keep[2].size, keep[1].size, keep[6].size - 1, *(a.ctypes.data for a in keep))
            capi.check(L.glu_plan_create(C.byref(pat), C.byref(opts), C.byref(h)), "glu_plan_create")
            self.analysis = None
        self._h = h
        info = np.zeros(12, dtype=np.int64)
        capi.check(L.glu_plan_info(h, capi.ptr(info)), "glu_plan_info")
        self.n, self.nnz_lu, self.nnz_a, self.num_lev, self.updates, self.segments, \
            self.device_bytes, self.units, self.lds_window_count, self.lds_window_steps, \
            self.lds_window_max_values, self.batch_grid_levels = (int(v) for v in info)

    @property
    def handle(self):
        return self._h

    # -- single matrix -------------------------------------------------
    def factor(self, lu: torch.Tensor, perturb: bool = False):
        """Factor FP32 values in the LU pattern in place (device)."""
        _check_dev(lu, torch.float32, self.nnz_lu, "lu")
        capi.check(capi.lib().glu_factor(self._h, _dptr(lu), int(bool(perturb)), _stream()), "glu_factor")
        return lu

    def factor_host(self, val: np.ndarray, perturb: bool = False):
        """LUonDevice contract: host FP32 values in, factors out, synchronous."""
        if val.dtype != np.float32 or not val.flags.c_contiguous or val.size != self.nnz_lu:
            raise ValueError("val must be a contiguous float32 array of nnz(L+U) entries")
        capi.check(capi.lib().glu_factor_host(self._h, capi.ptr(val), int(bool(perturb))), "glu_factor_host")
        return val

    def solve(self, lu: torch.Tensor, b: torch.Tensor, x: torch.Tensor = None, transpose=False):
        """FP32 solve (Symbolic_Matrix::solve semantics) on the device."""
        _check_dev(lu, torch.float32, self.nnz_lu, "lu")
        _check_dev(b, torch.float32, self.n, "b")
        if x is None:
            x = torch.empty_like(b)
        _check_dev(x, torch.float32, self.n, "x")
        capi.check(capi.lib().glu_solve(self._h, _dptr(lu), _dptr(b), _dptr(x), int(bool(transpose)),
                                        _stream()), "glu_solve")
        return x

    def refresh(self, ax: torch.Tensor, lu: torch.Tensor = None):
        """Scale + place FP64 value sets (batch x nnz_a) into FP32 LU buffers."""
        batch = ax.numel() // max(self.nnz_a, 1)
        _check_dev(ax, torch.float64, batch * self.nnz_a, "ax")
        if lu is None:
            lu = torch.empty(batch, self.nnz_lu, dtype=torch.float32, device=ax.device)
        _check_dev(lu, torch.float32, batch * self.nnz_lu, "lu")
        capi.check(capi.lib().glu_refresh(self._h, _dptr(ax), _dptr(lu), batch, _stream()), "glu_refresh")
        return lu

    def refine(self, ax, lu, b, x=None, tol=1e-10, maxit=10, transpose=False):
        """FP32-factor / FP64-residual iterative refinement.  Returns (x, iters, eta)."""
        _check_dev(ax, torch.float64, self.nnz_a, "ax")
        _check_dev(lu, torch.float32, self.nnz_lu, "lu")
        _check_dev(b, torch.float64, self.n, "b")
        if x is None:
            x = torch.empty_like(b)
        _check_dev(x, torch.float64, self.n, "x")
        it = C.c_int(0)
        eta = C.c_double(0.0)
        capi.check(capi.lib().glu_refine(self._h, _dptr(ax), _dptr(lu), _dptr(b), _dptr(x), float(tol),
                                         int(maxit), int(bool(transpose)), C.byref(it), C.byref(eta),
                                         _stream()), "glu_refine")
        return x, it.value, eta.value

    def adjoint_values(self, lam, x, grad=None):
        _check_dev(lam, torch.float64, self.n, "lam")
        _check_dev(x, torch.float64, self.n, "x")
        if grad is None:
            grad = torch.empty(self.nnz_a, dtype=torch.float64, device=x.device)
        _check_dev(grad, torch.float64, self.nnz_a, "grad")
        capi.check(capi.lib().glu_adjoint_values(self._h, _dptr(lam), _dptr(x), _dptr(grad), _stream()),
                   "glu_adjoint_values")
        return grad

    # -- same-pattern batch --------------------------------------------
    def batch_solve(self, ax, b, tol=1e-10, maxit=10, perturb=False, lu=None, x=None):
        """Refresh + factor + solve + refine of every matrix: a workgroup per
        matrix (after the wide leading steps as grid launches over the batch).

        ax: (B, nnz_a) float64, b: (B, n) float64.  Returns (x, lu, iters, eta).
        """
        B = b.shape[0]
        _check_dev(ax, torch.float64, B * self.nnz_a, "ax")
        _check_dev(b, torch.float64, B * self.n, "b")
        dev = b.device
        if x is None:
            x = torch.empty(B, self.n, dtype=torch.float64, device=dev)
        if lu is None:
            lu = torch.empty(B, self.nnz_lu, dtype=torch.float32, device=dev)
        _check_dev(x, torch.float64, B * self.n, "x")
        _check_dev(lu, torch.float32, B * self.nnz_lu, "lu")
        iters = torch.empty(B, dtype=torch.int32, device=dev)
        eta = torch.empty(B, dtype=torch.float64, device=dev)
        capi.check(capi.lib().glu_batch_solve(self._h, B, _dptr(ax), _dptr(b), _dptr(x), _dptr(lu), float(tol),
                                              int(maxit), _dptr(iters), _dptr(eta), int(bool(perturb)),
                                              _stream()), "glu_batch_solve")
        return x, lu, iters, eta

    def abft_sums(self, val):
        """ABFT checksums of values in the LU pattern before the factorization
        (Symbolic_Matrix::ABFTCalculateCCA): (B, nnz_lu) or (nnz_lu,) float32
        -> (B, n) or (n,) float32 column sums."""
        B = val.numel() // self.nnz_lu
        _check_dev(val, torch.float32, B * self.nnz_lu, "val")
        cca = torch.empty(B, self.n, dtype=torch.float32, device=val.device)
        capi.check(capi.lib().glu_abft_sums(self._h, B, _dptr(val), _dptr(cca), _stream()), "glu_abft_sums")
        return cca if val.dim() > 1 else cca.reshape(-1)

    def abft_check(self, lu, cca, tol=1e-5):
        """Symbolic_Matrix::ABFTCheckResult on factors lu against cca from
        abft_sums: (first failing column or -1, largest |difference|) per
        matrix, as int32 / float32 tensors."""
        B = lu.numel() // self.nnz_lu
        _check_dev(lu, torch.float32, B * self.nnz_lu, "lu")
        _check_dev(cca, torch.float32, B * self.n, "cca")
        first = torch.empty(B, dtype=torch.int32, device=lu.device)
        mx = torch.empty(B, dtype=torch.float32, device=lu.device)
        capi.check(capi.lib().glu_abft_check(self._h, B, _dptr(lu), _dptr(cca), float(tol), _dptr(first), _dptr(mx),
                                             _stream()), "glu_abft_check")
        return first, mx

    def batch_factor(self, ax, lu=None, perturb=False):
        """Refresh + factor of every matrix (no solve), launched as batch_solve's
        factorization.

        ax: (B, nnz_a) float64.  Returns lu (B, nnz_lu) float32, bit-identical
        to refresh + factor of each matrix.
        """
        B = ax.shape[0]
        _check_dev(ax, torch.float64, B * self.nnz_a, "ax")
        if lu is None:
            lu = torch.empty(B, self.nnz_lu, dtype=torch.float32, device=ax.device)
        _check_dev(lu, torch.float32, B * self.nnz_lu, "lu")
        capi.check(capi.lib().glu_batch_factor(self._h, B, _dptr(ax), _dptr(lu), int(bool(perturb)), _stream()),
                   "glu_batch_factor")
        return lu

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            capi.lib().glu_plan_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Group:
    """Distinct patterns refactored + refined together in one fused launch."""

    def __init__(self, plans, split=None):
        """split: how many of the heaviest members run on the single-matrix
        path beside the fused launch (0..2; None: the library's automatic
        choice, one in a group of four or more).  Same results either way."""
        self.plans = list(plans)
        arr = (C.c_void_p * len(self.plans))(*[p.handle.value for p in self.plans])
        h = C.c_void_p()
        capi.check(capi.lib().glu_group_create(arr, len(self.plans), C.byref(h)), "glu_group_create")
        self._h = h
        if split is not None:
            capi.check(capi.lib().glu_group_set_split(h, int(split)), "glu_group_set_split")

    def solve(self, axs, bs, tol=1e-10, maxit=10, perturb=False):
        cnt = len(self.plans)
        if len(axs) != cnt or len(bs) != cnt:
            raise ValueError("one value set and one rhs per plan")
        dev = bs[0].device
        xs, lus = [], []
        for p, ax, b in zip(self.plans, axs, bs):
            _check_dev(ax, torch.float64, p.nnz_a, "ax")
            _check_dev(b, torch.float64, p.n, "b")
            xs.append(torch.empty(p.n, dtype=torch.float64, device=dev))
            lus.append(torch.empty(p.nnz_lu, dtype=torch.float32, device=dev))
        arr = lambda ts: (C.c_void_p * cnt)(*[t.data_ptr() for t in ts])
        iters = torch.empty(cnt, dtype=torch.int32, device=dev)
        eta = torch.empty(cnt, dtype=torch.float64, device=dev)
        capi.check(capi.lib().glu_group_solve(self._h, arr(axs), arr(bs), arr(xs), arr(lus), float(tol),
                                              int(maxit), _dptr(iters), _dptr(eta), int(bool(perturb)),
                                              _stream()), "glu_group_solve")
        return xs, lus, iters, eta

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            capi.lib().glu_group_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _SpSolve(torch.autograd.Function):
    """x = A^-1 b with gradients for A's stored values and b.

    Forward: value refresh, FP32 factorization, FP64-refined solve.
    Backward: lam = A^-T g with the same FP32 factors (transposed solves)
    refined in FP64; dA = -lam x^T on A's pattern only; db = lam.
    """

    @staticmethod
    def forward(ctx, values, b, plan, tol, maxit):
        lu = plan.refresh(values.detach().contiguous()).reshape(-1)
        plan.factor(lu)
        x, it, eta = plan.refine(values.detach().contiguous(), lu, b.detach().contiguous(), tol=tol, maxit=maxit)
        ctx.plan, ctx.tol, ctx.maxit = plan, tol, maxit
        ctx.save_for_backward(values.detach(), lu, x)
        ctx.info = (it, eta)
        return x

    @staticmethod
    def backward(ctx, g):
        values, lu, x = ctx.saved_tensors
        plan = ctx.plan
        lam, _, _ = plan.refine(values.contiguous(), lu, g.contiguous().to(torch.float64), tol=ctx.tol,
                                maxit=ctx.maxit, transpose=True)
        gv = plan.adjoint_values(lam, x) if ctx.needs_input_grad[0] else None
        gb = lam if ctx.needs_input_grad[1] else None
        return gv, gb, None, None, None


def spsolve_autograd(values: torch.Tensor, b: torch.Tensor, plan: Plan, tol=1e-13, maxit=20):
    """Differentiable solve: ``values`` are A's stored entries (caller CSC
    order) as a float64 CUDA tensor; ``plan`` comes from an Analysis of the
    pattern."""
    if plan.analysis is None:
        raise ValueError("autograd needs a plan created from an Analysis")
    return _SpSolve.apply(values, b, plan, tol, maxit)
