"""GPU checks of the mixed-precision form: 16-bit storage, matrix-core
multi-RHS products, lockstep GMRES and iterative refinement.

* products: FP32 matrix-core product against numpy (1e-5); 16-bit products
  against numpy on the same rounded matrix (1e-5: only the FP32 accumulation
  order differs) and against the unrounded matrix within the format's
  precision (bf16 8 bits, fp16 11 bits of mantissa);
* solves: multi-RHS GMRES (complex64 and 16-bit) against a direct solve;
  refinement reaches 1e-6 on a BEM matrix and matches the FP64 oracle
  solution of the same discrete system.
"""
import numpy as np
import pytest
import torch

import oracle
from neuralat_amd import mesh

pytestmark = pytest.mark.gpu


def _dev():
    return torch.device("cuda:0")


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _rand(rng, *shape):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


def _round_planes(A, dtype, scale):
    t = torch.tensor(A * scale)
    if dtype == "bf16":
        r = t.real.to(torch.bfloat16).float().numpy()
        i = t.imag.to(torch.bfloat16).float().numpy()
    else:
        r = t.real.to(torch.float16).float().numpy()
        i = t.imag.to(torch.float16).float().numpy()
    return (r + 1j * i) / scale


# odd column counts take the guarded path, even ones the 16-byte-piece path
# with look-ahead (whole 256-column stages, then a guarded partial stage):
# 2048 columns are whole stages only, 1000 and 1538 end in a partial stage,
# 64 is a partial stage alone; 7 and 20 right-hand sides are ragged blocks
@pytest.mark.parametrize("split", [True, False])
@pytest.mark.parametrize("shape", [(100, 77, 3), (1000, 1000, 16), (257, 4097, 16), (64, 64, 20), (33, 50, 1),
                                   (512, 2048, 16), (300, 1538, 7)])
def test_fp32_matrix_core_product(shape, split):
    from neuralat_amd.bem.precision import matmul

    rng = np.random.default_rng(0)
    r, c, k = shape
    A, X, Y0 = _rand(rng, r, c), _rand(rng, k, c), _rand(rng, k, r)
    out = torch.tensor(Y0, device=_dev())
    matmul(torch.tensor(A, device=_dev()), torch.tensor(X, device=_dev()), alpha=0.5 - 1j, beta=2.0, out=out,
           split=split)
    ref = (0.5 - 1j) * (X.astype(np.complex128) @ A.T.astype(np.complex128)) + 2.0 * Y0
    assert _rel(out.cpu().numpy(), ref) < 1e-5, (shape, split)


@pytest.mark.parametrize("dtype", ["bf16", "fp16"])
@pytest.mark.parametrize("shape", [(100, 77, 3), (1000, 1000, 16), (257, 4097, 16), (64, 64, 20), (130, 70, 1)])
def test_lowp_products(dtype, shape):
    from neuralat_amd.bem.precision import LowPrecisionMatrix

    rng = np.random.default_rng(1)
    r, c, k = shape
    A, X = _rand(rng, r, c), _rand(rng, k, c)
    A[0, 0] *= 300.0  # exercise the fp16 scale
    L = LowPrecisionMatrix(torch.tensor(A, device=_dev()), dtype)
    Y = L.matmul(torch.tensor(X, device=_dev())).cpu().numpy()
    Ar = _round_planes(A, dtype, L.scale)
    # right-hand sides are split into two 16-bit parts (hi + lo) inside the
    # product, which leaves ~2^-16 (bf16) / 2^-22 (fp16) of rounding
    Xr = X
    tol_same = 1e-5 if (dtype == "fp16" or k == 1) else 4e-5
    assert _rel(Y, Xr.astype(np.complex128) @ Ar.T.astype(np.complex128)) < tol_same, (dtype, shape)
    bound = {"bf16": 1.2e-2, "fp16": 1.5e-3}[dtype]
    assert _rel(Y, X.astype(np.complex128) @ A.T.astype(np.complex128)) < bound


@pytest.mark.parametrize("kind", ["c64", "bf16", "fp16"])
def test_gmres_multi_vs_direct(kind):
    from neuralat_amd.bem.precision import LowPrecisionMatrix, gmres_multi

    rng = np.random.default_rng(2)
    n, k = 600, 7
    A = (np.eye(n) * 2 + _rand(rng, n, n) / np.sqrt(n) * 0.6).astype(np.complex64)
    B = _rand(rng, k, n)
    At = torch.tensor(A, device=_dev())
    op = At if kind == "c64" else LowPrecisionMatrix(At, kind)
    tol = 1e-6 if kind == "c64" else 1e-4
    X, info = gmres_multi(op, torch.tensor(B, device=_dev()), tol=tol, restart=50, maxiter=400)
    assert all(i["converged"] for i in info), info
    Aop = A if kind == "c64" else _round_planes(A, kind, op.scale)
    ref = np.linalg.solve(Aop.astype(np.complex128), B.T.astype(np.complex128)).T
    assert _rel(X.cpu().numpy(), ref) < 10 * tol


@pytest.mark.parametrize("dtype", ["bf16", "fp16"])
def test_iterative_refinement_on_bem_system(dtype):
    from neuralat_amd.bem import BEM_Solver
    from neuralat_amd.bem.precision import LowPrecisionMatrix, solve_ir

    v, t = mesh.icosphere(3, 1.0)
    v = mesh.sh_perturb(v, 4, 0.05, seed=3).astype(np.float32)
    k = 8.0
    s = BEM_Solver(torch.tensor(v, device=_dev()), torch.tensor(t, device=_dev()), basis="DP0")
    rng = np.random.default_rng(5)
    srcs = rng.uniform(-0.3, 0.3, size=(16, 3))
    psis = []
    for q in range(16):
        _, dn = mesh.sources_field([srcs[q]], None, k)
        psis.append(mesh.project_dp0(v, t, lambda x, nn: dn(x, nn)))
    psi = torch.tensor(np.array(psis, dtype=np.complex64), device=_dev())
    # one fused pass per right-hand side gives the same lhs; keep the last
    rhs = []
    for q in range(16):
        lhs, r = s.cbie_system(k, psi[q])
        rhs.append(r)
    B = torch.stack(rhs)
    L = LowPrecisionMatrix(lhs, dtype)
    X, info = solve_ir(lhs, L, B, tol=1e-6, inner_tol=1e-3 if dtype == "fp16" else 1e-2)
    assert info["converged"] and max(info["rel_residual"]) <= 1e-6, info
    # FP64 oracle of the same discrete system, two of the sixteen systems
    S = oracle.assemble(v, t, k, "single", "DP0")
    D = oracle.assemble(v, t, k, "double", "DP0")
    M = oracle.identity(v, t, "DP0")
    A64 = 0.5 * M - D
    for q in (0, 11):
        x64 = np.linalg.solve(A64, -S @ psis[q])
        assert _rel(X[q].cpu().numpy(), x64) < 1e-4, q
