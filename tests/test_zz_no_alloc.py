"""GMRES without allocation (bem_gmres_noalloc) against the same solve with a
caller's workspace (bem_gmres_ex), both through the C ABI: two calls in a
row, the second on a smaller system (and on A^H), give the same iterates bit
for bit and the solution of a direct FP64 solve; a size beyond the static
scratch is refused before any GPU work."""
import ctypes

import numpy as np
import pytest
import torch

from neuralat_amd import _lib

pytestmark = pytest.mark.gpu


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _system(n, seed):
    rng = np.random.default_rng(seed)
    A = (np.eye(n) * (1.5 + 0.5j) + (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
         * (0.9 / np.sqrt(n))).astype(np.complex64)
    b = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    return A, b


def _solve(lib, At, bt, trans, restart, noalloc):
    n = At.shape[0]
    x = torch.zeros(n, dtype=torch.complex64, device="cuda")
    info = _lib.SolveInfo()
    if noalloc:
        rc = lib.bem_gmres_noalloc(_ptr(At), n, n, trans, _ptr(bt), _ptr(x), 1e-6, restart, 2000,
                                   ctypes.byref(info), _stream())
    else:
        work = torch.empty(lib.bem_gmres_work_size(n, restart), dtype=torch.uint8, device="cuda")
        rc = lib.bem_gmres_ex(_ptr(At), n, n, trans, _ptr(bt), _ptr(x), 1e-6, restart, 2000, _ptr(work),
                              ctypes.byref(info), _stream())
    torch.cuda.synchronize()
    return rc, x, info.as_dict()


@pytest.mark.parametrize("order", ["large_then_small"])
def test_gmres_noalloc_matches_workspace_solve(order):
    lib = _lib.load()
    cases = [(5000, 60, 0, 11), (1000, 40, 2, 12)]  # (n, restart, trans, seed): multi-kernel step, then fused
    for n, restart, trans, seed in cases:
        print(f"bem_gmres_noalloc: n = {n}, restart = {restart}, trans = {trans}", flush=True)
        assert lib.bem_gmres_work_size(n, restart) <= lib.bem_gmres_noalloc_capacity()
        A, b = _system(n, seed)
        At, bt = torch.tensor(A, device="cuda"), torch.tensor(b, device="cuda")
        rc0, x0, i0 = _solve(lib, At, bt, trans, restart, noalloc=False)
        rc1, x1, i1 = _solve(lib, At, bt, trans, restart, noalloc=True)
        assert rc0 == _lib.BEM_OK and rc1 == _lib.BEM_OK, (rc0, rc1, _lib.last_error())
        assert i0 == i1 and i1["converged"], (i0, i1)
        assert torch.equal(x0, x1)
        Aop = A.astype(np.complex128)
        Aop = {0: Aop, 2: Aop.conj().T}[trans]
        ref = np.linalg.solve(Aop, b.astype(np.complex128))
        assert np.linalg.norm(x1.cpu().numpy() - ref) / np.linalg.norm(ref) < 1e-5


def test_gmres_noalloc_refuses_oversize():
    lib = _lib.load()
    n = 20480
    restart = 1
    while lib.bem_gmres_work_size(n, restart + 1) <= lib.bem_gmres_noalloc_capacity():
        restart += 1
    # refused on the host, before any device access: the pointers are never read
    A = torch.zeros(16, dtype=torch.complex64, device="cuda")
    info = _lib.SolveInfo()
    rc = lib.bem_gmres_noalloc(_ptr(A), n, n, 0, _ptr(A), _ptr(A), 1e-6, restart + 1, 10, ctypes.byref(info),
                               _stream())
    assert rc == _lib.BEM_ERR_BAD_ARG and "static scratch" in _lib.last_error()
    with pytest.raises(ValueError):
        _lib.check(rc, "bem_gmres_noalloc")
