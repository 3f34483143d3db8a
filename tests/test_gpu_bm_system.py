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


def test_fused_systems_over_wavenumbers_and_samples():
    """The batched fused systems (bem_burton_miller_systems over wavenumbers,
    bem_burton_miller_samples over a geometry set) equal the one-system calls
    and the FP64 oracle."""
    from neuralat_amd.bem import BEM_Solver, GeometrySet, Mesh, burton_miller_samples

    dev = torch.device("cuda:0")
    v, t = mesh.icosphere(3, 0.1)
    v = mesh.sh_perturb(v, 6, 0.05, seed=9).astype(np.float32)
    rng = np.random.default_rng(5)
    ks = [5.0, 12.0, -7.5]
    G = (rng.standard_normal((3, len(v))) + 1j * rng.standard_normal((3, len(v)))).astype(np.complex64)
    s = BEM_Solver(torch.tensor(v, device=dev), torch.tensor(t, device=dev), basis="P1", clamp=False)
    print("fused systems over wavenumbers", flush=True)
    lb, rb = s.burton_miller_systems(ks, torch.tensor(G, device=dev))
    for z, k in enumerate(ks):
        l1, r1 = s.burton_miller_system(k, torch.tensor(G[z], device=dev))
        assert _rel(lb[z].cpu().numpy(), l1.cpu().numpy()) < 1e-6, z
        assert _rel(rb[z].cpu().numpy(), r1.cpu().numpy()) < 1e-6, z
    M = oracle.identity(v, t, "P1")
    R64 = oracle.assemble(v, t, ks[1], "bm_rhs", "P1", quad="approx", clamp=False) - 0.5j / ks[1] * M
    assert _rel(rb[1].cpu().numpy(), R64 @ G[1].astype(np.complex128)) < 1e-5
    # three placements of the mesh (rotated, scaled, moved), each at its own k
    vs = []
    for q in range(3):
        th = 0.4 * q
        rot = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
        vs.append(((v @ rot.T) * (1.0 + 0.2 * q) + np.array([0.05 * q, 0, 0])).astype(np.float32))
    m = Mesh(torch.tensor(vs[0], device=dev), torch.tensor(t, device=dev))
    gs = GeometrySet(m, torch.tensor(np.stack(vs), device=dev))
    print("fused systems over geometry-set samples", flush=True)
    ls, rs = burton_miller_samples(gs, ks, torch.tensor(G, device=dev), clamp=False)
    for z in range(3):
        sz = BEM_Solver(torch.tensor(vs[z], device=dev), torch.tensor(t, device=dev), basis="P1", clamp=False)
        l1, r1 = sz.burton_miller_system(ks[z], torch.tensor(G[z], device=dev))
        assert _rel(ls[z].cpu().numpy(), l1.cpu().numpy()) < 1e-6, z
        assert _rel(rs[z].cpu().numpy(), r1.cpu().numpy()) < 1e-6, z
    M2 = oracle.identity(vs[2], t, "P1")
    R64 = oracle.assemble(vs[2], t, ks[2], "bm_rhs", "P1", quad="approx", clamp=False) - 0.5j / ks[2] * M2
    assert _rel(rs[2].cpu().numpy(), R64 @ G[2].astype(np.complex128)) < 1e-5
