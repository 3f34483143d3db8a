"""Batched (mesh, k) items: one persistent assembly launch for many DP0 CBIE
systems of uneven size, and lockstep graph-captured GMRES over them.

No reference counterpart: the reference's data generation solves one item at
a time (src/scene.py:270-291, experiments/neuPAT/generate.py:23-33).
"""
from __future__ import annotations

import ctypes

import torch

from .. import _lib
from .solver import Mesh, _ptr, _stream, check_tensor


class _Item(ctypes.Structure):
    _fields_ = [("mesh", ctypes.c_void_p), ("k", ctypes.c_float), ("psi", ctypes.c_void_p),
                ("lhs", ctypes.c_void_p), ("ld", ctypes.c_int64), ("rhs", ctypes.c_void_p)]


class _Sys(ctypes.Structure):
    _fields_ = [("A", ctypes.c_void_p), ("lda", ctypes.c_int64), ("n", ctypes.c_int64), ("b", ctypes.c_void_p),
                ("x", ctypes.c_void_p)]


def _check(rc, what):
    if rc != _lib.BEM_OK:
        msg = _lib.load().bem_batch_last_error().decode()
        if rc == _lib.BEM_ERR_NOT_CONVERGED:
            raise _lib.NotConvergedError(f"{what}: {msg}")
        if rc == _lib.BEM_ERR_BAD_ARG:
            raise ValueError(f"{what}: {msg or 'bad argument'}")
        raise _lib.BemError(f"{what}: {msg} (status {rc})")


def _items(meshes, ks, psis, outs):
    arr = (_Item * len(meshes))()
    for q, (m, k, psi, (lhs, rhs)) in enumerate(zip(meshes, ks, psis, outs)):
        check_tensor(psi, torch.complex64)
        arr[q] = _Item(m.handle.value, float(k), psi.data_ptr(), lhs.data_ptr(), lhs.stride(0), rhs.data_ptr())
    return arr


def allocate_systems(meshes, device):
    """(lhs [nt,nt], rhs [nt]) complex64 buffers for each mesh."""
    return [(torch.empty((m.nt, m.nt), dtype=torch.complex64, device=device),
             torch.empty(m.nt, dtype=torch.complex64, device=device)) for m in meshes]


def batch_cbie(meshes, ks, psis, outs=None, clamp=True, row_tile=0):
    """Fused DP0 CBIE systems (M/2 - K, -V psi) for items (meshes[q], ks[q],
    psis[q]) in one persistent launch.  Returns [(lhs, rhs), ...]."""
    device = psis[0].device
    outs = outs or allocate_systems(meshes, device)
    arr = _items(meshes, ks, psis, outs)
    lib = _lib.load()
    work = torch.empty(lib.bem_batch_cbie_work_size(arr, len(arr)), dtype=torch.uint8, device=device)
    _check(lib.bem_batch_cbie(arr, len(arr), int(clamp), int(row_tile), _ptr(work), _stream()), "bem_batch_cbie")
    return outs


def batch_autotune(meshes, ks, psis, outs=None, clamp=True):
    """Pick the rows-per-task tile of the persistent kernel by timing on the
    largest item; the choice is kept for its size class.  Returns
    (best, {tile: ms})."""
    device = psis[0].device
    outs = outs or allocate_systems(meshes, device)
    arr = _items(meshes, ks, psis, outs)
    lib = _lib.load()
    work = torch.empty(lib.bem_batch_cbie_work_size(arr, len(arr)), dtype=torch.uint8, device=device)
    best = ctypes.c_int()
    ms = (ctypes.c_float * 3)()
    _check(lib.bem_batch_autotune(arr, len(arr), int(clamp), _ptr(work), ctypes.byref(best), ms, _stream()),
           "bem_batch_autotune")
    return int(best.value), dict(zip((16, 32, 64), list(ms)))


def _zero_vectors(bs):
    """Zero initial guesses shaped like ``bs``: views of one allocation (one
    fill for the whole batch)."""
    flat = torch.zeros(sum(b.numel() for b in bs), dtype=bs[0].dtype, device=bs[0].device)
    return [v.view(b.shape) for v, b in zip(flat.split([b.numel() for b in bs]), bs)]


def batch_gmres(As, bs, xs=None, tol=1e-6, restart=50, max_cycles=40, use_graph=True, raise_on_fail=False):
    """Lockstep GMRES on systems As[q] x = bs[q] (different sizes).  Returns
    (xs, infos)."""
    n = len(As)
    xs = xs or _zero_vectors(bs)
    arr = (_Sys * n)()
    for q in range(n):
        check_tensor(As[q], torch.complex64)
        check_tensor(bs[q], torch.complex64)
        arr[q] = _Sys(As[q].data_ptr(), As[q].stride(0), As[q].shape[0], bs[q].data_ptr(), xs[q].data_ptr())
    lib = _lib.load()
    restart = int(min(restart, max(A.shape[0] for A in As)))
    work = torch.empty(lib.bem_batch_gmres_work_size(arr, n, restart), dtype=torch.uint8, device=bs[0].device)
    info = (_lib.SolveInfo * n)()
    rc = lib.bem_batch_gmres(arr, n, float(tol), restart, int(max_cycles), int(use_graph), _ptr(work), info,
                             _stream())
    if rc != _lib.BEM_ERR_NOT_CONVERGED or raise_on_fail:
        _check(rc, "bem_batch_gmres")
    return xs, [i.as_dict() for i in info]


def batch_solve_mixed(As, bs, xs=None, dtype="bf16", tol=1e-6, inner_tol=1e-6, outer_restart=10,
                      inner_restart=200, max_inner_cycles=10, max_outer_cycles=2, use_graph=True,
                      raise_on_fail=False):
    """Lockstep mixed-precision solves of As[q] x = bs[q]: flexible GMRES in
    FP32 on the complex64 matrices, preconditioned by lockstep GMRES on 16-bit
    (``dtype`` "bf16" or "fp16") or 8-bit ("fp8": OCP e4m3, a power-of-two
    scale per row) copies with the diagonal kept in FP32.
    Returns (xs, infos); infos[q]["iterations"] = outer steps,
    infos[q]["inner_iterations"] = low-precision Arnoldi steps."""
    n = len(As)
    codes = {"bf16": 1, "fp16": 2, "fp8": 3}
    if dtype not in codes:
        raise ValueError("dtype must be 'bf16', 'fp16' or 'fp8'")
    xs = xs or _zero_vectors(bs)
    arr = (_Sys * n)()
    for q in range(n):
        check_tensor(As[q], torch.complex64)
        check_tensor(bs[q], torch.complex64)
        arr[q] = _Sys(As[q].data_ptr(), As[q].stride(0), As[q].shape[0], bs[q].data_ptr(), xs[q].data_ptr())
    lib = _lib.load()
    inner_restart = int(min(inner_restart, max(A.shape[0] for A in As)))
    outer_restart = int(outer_restart)
    work = torch.empty(lib.bem_batch_fgmres_work_size(arr, n, outer_restart, inner_restart), dtype=torch.uint8,
                       device=bs[0].device)
    info = (_lib.SolveInfo * n)()
    inner = (ctypes.c_int32 * n)()
    rc = lib.bem_batch_fgmres(arr, n, codes[dtype], float(tol), float(inner_tol), outer_restart,
                              inner_restart, int(max_inner_cycles), int(max_outer_cycles), int(use_graph),
                              _ptr(work), info, inner, _stream())
    if rc != _lib.BEM_ERR_NOT_CONVERGED or raise_on_fail:
        _check(rc, "bem_batch_fgmres")
    out = []
    for q in range(n):
        d = info[q].as_dict()
        d["inner_iterations"] = int(inner[q])
        out.append(d)
    return xs, out


__all__ = ["Mesh", "allocate_systems", "batch_cbie", "batch_autotune", "batch_gmres", "batch_solve_mixed"]
