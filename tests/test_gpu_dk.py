"""Wavenumber derivatives of the P1 operators and of the potential matrices
(bem_assemble_dk, bem_potential_dk) against central differences of the FP64
oracle in k.

The reference has no derivatives; the oracle's operators are pinned by the
golden fixtures and the invariant tests, and a central difference with
h = 1e-4 |k| in FP64 has a truncation error near 1e-9 relative (the kernels
are analytic in k) and a rounding error near 1e-12, far under the FP32 bar.
Bar: relative Frobenius error <= 1e-5, the plain operators' parity bar.
"""
import numpy as np
import pytest
import torch

import oracle
from neuralat_amd import mesh

pytestmark = pytest.mark.gpu

OPS = ["single", "double", "adjointdouble", "bm_lhs", "bm_rhs"]
TOL = 1e-5


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _solver(v, t, basis="P1", clamp=True):
    from neuralat_amd.bem import BEM_Solver

    return BEM_Solver(torch.tensor(np.asarray(v, np.float32), device="cuda:0"),
                      torch.tensor(np.asarray(t, np.int32), device="cuda:0"), basis=basis, clamp=clamp)


def _fd_operator(v, t, k, op, quad, clamp):
    h = 1e-4 * abs(k)
    hi = oracle.assemble(v, t, k + h, op, "P1", quad=quad, clamp=clamp)
    lo = oracle.assemble(v, t, k - h, op, "P1", quad=quad, clamp=clamp)
    return (hi - lo) / (2 * h)


@pytest.fixture(scope="module")
def l2_sphere():
    v, t = mesh.icosphere(2, 1.0)
    v = mesh.sh_perturb(v, 4, 0.05, seed=2)
    return v.astype(np.float32), t


@pytest.mark.parametrize("quad", ["full", "approx"])
@pytest.mark.parametrize("op", OPS)
def test_operator_dk_vs_oracle_difference(l2_sphere, op, quad):
    v, t = l2_sphere
    s = _solver(v, t)
    for k in (4.0, -9.5):
        print("dA/dk", op, quad, k, flush=True)
        A = s.assemble_boundary_matrix_dk(k, op, approx=quad == "approx").cpu().numpy()
        e = _rel(A, _fd_operator(v, t, k, op, quad, True))
        assert e < TOL, (op, quad, k, e)


@pytest.mark.parametrize("op", ["bm_lhs", "bm_rhs"])
def test_operator_dk_many_tiles_no_clamp(op):
    # level 3: five column tiles, vertices shared between tiles
    v, t = mesh.icosphere(3, 0.1)
    v = mesh.sh_perturb(v, 6, 0.05, seed=4).astype(np.float32)
    s = _solver(v, t, clamp=False)
    k = 18.3
    print("dA/dk level 3", op, flush=True)
    A = s.assemble_boundary_matrix_dk(k, op, approx=True).cpu().numpy()
    e = _rel(A, _fd_operator(v, t, k, op, "approx", False))
    assert e < TOL, (op, e)


@pytest.mark.parametrize("basis", ["P1", "DP0"])
@pytest.mark.parametrize("op", ["single", "double"])
def test_potential_dk_vs_oracle_difference(l2_sphere, op, basis):
    v, t = l2_sphere
    s = _solver(v, t, basis)
    rng = np.random.default_rng(3)
    d = rng.standard_normal((40, 3))
    pts = (d / np.linalg.norm(d, axis=1, keepdims=True) * rng.uniform(1.5, 3.0, (40, 1))).astype(np.float32)
    for k in (6.0, -14.0):
        print("dP/dk", op, basis, k, flush=True)
        P = s.potential_matrix_dk(k, torch.tensor(pts, device="cuda:0"), op).cpu().numpy()
        h = 1e-4 * abs(k)
        ref = (oracle.potential(v, t, pts, k + h, op, basis) - oracle.potential(v, t, pts, k - h, op, basis)) / (2 * h)
        e = _rel(P, ref)
        assert e < TOL, (op, basis, k, e)


def test_dk_rejects_hypersingular_and_dp0(l2_sphere):
    from neuralat_amd import _lib
    from neuralat_amd.bem.solver import _ptr, _stream

    v, t = l2_sphere
    s = _solver(v, t)
    out = torch.empty((len(v), len(v)), dtype=torch.complex64, device="cuda:0")
    lib = _lib.load()
    assert lib.bem_assemble_dk(s.mesh.handle, 3.0, _lib.OPS["hypersingular"], 0, 1, _ptr(out), len(v),
                               _stream()) == _lib.BEM_ERR_BAD_ARG
    assert lib.bem_assemble_dk(s.mesh.handle, 0.0, _lib.OPS["bm_lhs"], 1, 1, _ptr(out), len(v),
                               _stream()) == _lib.BEM_ERR_BAD_ARG
    s0 = _solver(v, t, "DP0")
    with pytest.raises(ValueError):
        s0.assemble_boundary_matrix_dk(3.0, "single")
