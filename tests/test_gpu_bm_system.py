"""The fused Burton-Miller system (bem_burton_miller_system:
lhs = bm_lhs + M/2, rhs = (bm_rhs - (i/2k) M) g with bm_rhs contracted inside
the LHS kernels, never stored) against the FP64 oracle's assembled
operators (src/bem/solver.py:118-126 built from bind.cu:192-215) and
against the unfused route through the library (assemble bm_rhs, add the mass
term, one GEMV).  Bar: 1e-5 relative against FP64 (the operators' parity
bar), 1e-6 against the unfused FP32 route."""
import numpy as np
import pytest
import torch

import oracle
from neuralat_amd import mesh

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("clamp", [False, True])
@pytest.mark.parametrize("quad", ["approx", "full"])
@pytest.mark.parametrize("k", [18.3, -7.0])
def test_fused_burton_miller_system(k, quad, clamp):
    from neuralat_amd.bem import BEM_Solver
    from neuralat_amd.bem.solver import add_identity_, matvec

    # level 3 (642 vertices, five column tiles), radius 0.1
    v, t = mesh.icosphere(3, 0.1)
    v = mesh.sh_perturb(v, 6, 0.05, seed=8).astype(np.float32)
    rng = np.random.default_rng(4)
    g = (rng.standard_normal(len(v)) + 1j * rng.standard_normal(len(v))).astype(np.complex64)
    dev = torch.device("cuda:0")
    s = BEM_Solver(torch.tensor(v, device=dev), torch.tensor(t, device=dev), basis="P1", clamp=clamp)
    gt = torch.tensor(g, device=dev)
    print(f"fused Burton-Miller system k={k} {quad} clamp={clamp}", flush=True)
    lhs, rhs = s.burton_miller_system(k, gt, approx=quad == "approx")
    lhs, rhs = lhs.cpu().numpy(), rhs.cpu().numpy()
    M = oracle.identity(v, t, "P1")
    L64 = oracle.assemble(v, t, k, "bm_lhs", "P1", quad=quad, clamp=clamp) + 0.5 * M
    R64 = oracle.assemble(v, t, k, "bm_rhs", "P1", quad=quad, clamp=clamp) - 0.5j / k * M
    e_l, e_r = _rel(lhs, L64), _rel(rhs, R64 @ g.astype(np.complex128))
    R = add_identity_(s, s.assemble_boundary_matrix(k, "bm_rhs", approx=quad == "approx"), -0.5j / k)
    r32 = matvec(R, gt).cpu().numpy()
    e_u = _rel(rhs, r32)
    print(f"  lhs vs FP64 {e_l:.2e}, rhs vs FP64 {e_r:.2e}, rhs vs unfused {e_u:.2e}", flush=True)
    assert e_l < 1e-5 and e_r < 1e-5, (e_l, e_r)
    assert e_u < 1e-6, e_u
