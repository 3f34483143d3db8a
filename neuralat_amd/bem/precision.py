"""Mixed-precision solves: 16-bit matrix storage, matrix-core multi-RHS
products with FP32 accumulation, FP32 iterative refinement.

No reference counterpart (the reference assembles and solves in FP32 only,
src/bem/solver.py:61-104); the solution is checked against the FP32 path and
the FP64 oracle in tests/test_gpu_precision.py.
"""
from __future__ import annotations

import ctypes

import torch

from .. import _lib
from .solver import _ptr, _stream, check_tensor

DTYPES = {"bf16": 1, "fp16": 2}


class LowPrecisionMatrix:
    """Two 16-bit planes holding round(scale * A) for a complex64 matrix A,
    in the library's tiled matrix-core layout."""

    def __init__(self, A: torch.Tensor, dtype: str = "bf16", scale: float | None = None):
        check_tensor(A, torch.complex64)
        if dtype not in DTYPES:
            raise ValueError("dtype must be 'bf16' or 'fp16'")
        lib = _lib.load()
        self.rows, self.cols = A.shape
        self.dtype = dtype
        self.ld = int(lib.bem_lowp_ld(self.cols))
        # tiled for the matrix cores (16 x 32 tiles), see bem_convert_lowp
        self.planes = torch.empty((2, int(lib.bem_lowp_plane_elems(self.rows, self.cols))), dtype=torch.int16,
                                  device=A.device)
        used = ctypes.c_float()
        _lib.check(lib.bem_convert_lowp(_ptr(A), self.rows, self.cols, A.stride(0), DTYPES[dtype],
                                        float(scale or 0.0), _ptr(self.planes[0]), _ptr(self.planes[1]), self.ld,
                                        ctypes.byref(used), _stream()), "bem_convert_lowp")
        self.scale = float(used.value)
        self.device = A.device

    @property
    def Ar(self):
        return self.planes[0]

    @property
    def Ai(self):
        return self.planes[1]

    def matmul(self, X: torch.Tensor, alpha=1.0, beta=0.0, out=None):
        """Y[s] = alpha A X[s] + beta Y[s] for X of shape [nrhs, cols]."""
        X = _as_block(X)
        nrhs = X.shape[0]
        if out is None:
            out = torch.empty((nrhs, self.rows), dtype=torch.complex64, device=self.device)
        lib = _lib.load()
        work = torch.empty(lib.bem_matmul_lowp_work_size(self.ld), dtype=torch.uint8, device=self.device)
        a, b = complex(alpha), complex(beta)
        _lib.check(lib.bem_matmul_lowp(_ptr(self.Ar), _ptr(self.Ai), DTYPES[self.dtype], self.rows, self.cols, self.ld,
                                       1.0 / self.scale, _ptr(X), X.stride(0), nrhs, _ptr(out), out.stride(0), a.real,
                                       a.imag, b.real, b.imag, _ptr(work), _stream()), "bem_matmul_lowp")
        return out


def _as_block(X):
    if X.dim() == 1:
        X = X.unsqueeze(0)
    X = X.to(torch.complex64).contiguous()
    return X


def matmul(A: torch.Tensor, X: torch.Tensor, alpha=1.0, beta=0.0, out=None, split=True):
    """Y[s] = alpha A X[s] + beta Y[s] on the FP32 matrix cores (X: [nrhs, cols]).
    ``split``: give the product the scratch of a column split (used when the
    row tiles alone do not fill the GPU; the result is the same either way)."""
    check_tensor(A, torch.complex64)
    X = _as_block(X)
    nrhs = X.shape[0]
    rows, cols = A.shape
    if out is None:
        out = torch.empty((nrhs, rows), dtype=torch.complex64, device=A.device)
    a, b = complex(alpha), complex(beta)
    lib = _lib.load()
    work = torch.empty(lib.bem_matmul_work_size(rows), dtype=torch.uint8, device=A.device) if split else None
    _lib.check(lib.bem_matmul(_ptr(A), rows, cols, A.stride(0), _ptr(X), X.stride(0), nrhs, _ptr(out),
                              out.stride(0), a.real, a.imag, b.real, b.imag, _ptr(work) if split else None,
                              _stream()), "bem_matmul")
    return out


def _op_args(A):
    if isinstance(A, LowPrecisionMatrix):
        return None, 0, A.Ar, A.Ai, DTYPES[A.dtype], A.ld, 1.0 / A.scale, A.rows, A.device
    check_tensor(A, torch.complex64)
    return A, A.stride(0), None, None, 0, 0, 1.0, A.shape[0], A.device


def gmres_multi(A, B: torch.Tensor, tol=1e-6, restart=100, maxiter=1000, raise_on_fail=False):
    """Lockstep GMRES on the systems A x_s = B[s] (A complex64 or a
    LowPrecisionMatrix).  Returns (X [nrhs, n], list of per-system info)."""
    Af, lda, Ar, Ai, dt, ldlp, inv, n, dev = _op_args(A)
    B = _as_block(B)
    nrhs = B.shape[0]
    X = torch.zeros((nrhs, n), dtype=torch.complex64, device=dev)
    lib = _lib.load()
    restart = int(min(restart, n))
    work = torch.empty(lib.bem_gmres_multi_work_size(n, nrhs, restart, ldlp), dtype=torch.uint8, device=dev)
    info = (_lib.SolveInfo * nrhs)()
    rc = lib.bem_gmres_multi(_ptr(Af) if Af is not None else None, lda, _ptr(Ar) if Ar is not None else None,
                             _ptr(Ai) if Ai is not None else None, dt, ldlp, inv, n, nrhs, _ptr(B), B.stride(0),
                             _ptr(X), X.stride(0), float(tol), restart, int(maxiter), _ptr(work), info, _stream())
    if rc != _lib.BEM_ERR_NOT_CONVERGED or raise_on_fail:
        _lib.check(rc, "bem_gmres_multi")
    return X, [i.as_dict() for i in info]


def solve_ir(A: torch.Tensor, Alp: LowPrecisionMatrix, B: torch.Tensor, tol=1e-6, inner_tol=1e-3, restart=100,
             max_inner=1000, max_outer=30, raise_on_fail=False):
    """Mixed-precision refinement: flexible GMRES in FP32 on A (FP32 residuals
    and updates) whose directions are preconditioned by inner lockstep GMRES
    solves on the 16-bit Alp to inner_tol.  Returns (X, info) with info =
    {"outer", "inner", "rel_residual": [...], "converged"}."""
    check_tensor(A, torch.complex64)
    B = _as_block(B)
    nrhs, n = B.shape
    X = torch.zeros((nrhs, n), dtype=torch.complex64, device=A.device)
    lib = _lib.load()
    restart = int(min(restart, n))
    work = torch.empty(lib.bem_solve_ir_work_size(n, nrhs, restart, Alp.ld), dtype=torch.uint8, device=A.device)
    info = (_lib.SolveInfo * nrhs)()
    outer = ctypes.c_int32()
    rc = lib.bem_solve_ir(_ptr(A), A.stride(0), _ptr(Alp.Ar), _ptr(Alp.Ai), DTYPES[Alp.dtype], Alp.ld,
                          1.0 / Alp.scale, n, nrhs, _ptr(B), B.stride(0), _ptr(X), X.stride(0), float(tol),
                          float(inner_tol), restart, int(max_inner), int(max_outer), _ptr(work), info,
                          ctypes.byref(outer), _stream())
    if rc != _lib.BEM_ERR_NOT_CONVERGED or raise_on_fail:
        _lib.check(rc, "bem_solve_ir")
    res = [i.rel_residual for i in info]
    return X, {"outer": int(outer.value), "inner": int(info[0].iterations), "rel_residual": res,
               "converged": all(i.converged for i in info)}
