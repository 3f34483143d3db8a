"""Gradients of the P1 Burton-Miller pressure (grad.burton_miller_pressure,
the reference's neumann2dirichlet + boundary2potential,
src/bem/solver.py:118-126,140-147) against the FP64 oracle pipeline.

The reference has no gradients.  The FP64 references are differences of the
oracle's own pipeline: in k a central difference with h = 1e-4 |k|
(truncation near 1e-9 relative); along complex directions of the Neumann
data a central difference, exact up to rounding because the loss is
quadratic in the data.  Loss: sum |p|^2 over 32 exterior listeners.
Bar: 1e-4 relative (FP32 operators and GMRES to 1e-6 against FP64; measured
1e-6 in k and 3e-7 along the Neumann directions).
"""
import numpy as np
import pytest
import torch

import oracle
from neuralat_amd import mesh

pytestmark = pytest.mark.gpu


def _case(k):
    v, t = mesh.icosphere(3, 0.1)
    v = mesh.sh_perturb(v, 6, 0.05, seed=5).astype(np.float32)
    src = np.array([[0.01, -0.005, 0.002]])
    _, dn = mesh.sources_field(src, None, k)
    psi = mesh.project_p1(v, t, lambda x, n: dn(x, n))
    pts = mesh.random_exterior_points(32, 0.2, 0.5, seed=1)
    return v, t, psi, pts


def _fp64_ops(v, t, k, pts):
    M = oracle.identity(v, t, "P1")
    L = oracle.assemble(v, t, k, "bm_lhs", "P1", quad="approx", clamp=False) + 0.5 * M
    R = oracle.assemble(v, t, k, "bm_rhs", "P1", quad="approx", clamp=False) - 0.5j / k * M
    S = oracle.potential(v, t, pts, k, "single", "P1", clamp=False)
    D = oracle.potential(v, t, pts, k, "double", "P1", clamp=False)
    return L, R, S, D


def _loss64(ops, psi):
    L, R, S, D = ops
    p = D @ np.linalg.solve(L, R @ psi) - S @ psi
    return float(np.sum(np.abs(p) ** 2))


def _gpu(v, t, k, psi, pts, want_k=True):
    from neuralat_amd.bem.grad import burton_miller_pressure

    dev = torch.device("cuda:0")
    kt = torch.tensor(float(k), device=dev, requires_grad=want_k)
    g = torch.tensor(psi.astype(np.complex64), device=dev, requires_grad=True)
    p = burton_miller_pressure(torch.tensor(v, device=dev), torch.tensor(t, device=dev), kt, g,
                               torch.tensor(pts.astype(np.float32), device=dev), tol=1e-6)
    loss = (p.abs() ** 2).sum()
    loss.backward()
    return float(loss), (float(kt.grad) if want_k else None), g.grad.cpu().numpy()


@pytest.mark.parametrize("k", [18.3, -18.3, 7.0])
def test_wavenumber_gradient_vs_fp64_difference(k):
    v, t, psi, pts = _case(k)
    print(f"BM pressure gradient, k = {k}: GPU", flush=True)
    loss, gk, _ = _gpu(v, t, k, psi, pts)
    print(f"BM pressure gradient, k = {k}: FP64 oracle differences", flush=True)
    h = 1e-4 * abs(k)
    ref = (_loss64(_fp64_ops(v, t, k + h, pts), psi) - _loss64(_fp64_ops(v, t, k - h, pts), psi)) / (2 * h)
    l64 = _loss64(_fp64_ops(v, t, k, pts), psi)
    print(f"  loss {loss:.6e} (FP64 {l64:.6e}); dl/dk {gk:.6e} (FP64 {ref:.6e})", flush=True)
    assert abs(loss - l64) / l64 < 1e-4
    assert abs(gk - ref) / abs(ref) < 1e-4, (gk, ref)


@pytest.mark.parametrize("k", [18.3, -7.0])
def test_neumann_gradient_vs_fp64_difference(k):
    v, t, psi, pts = _case(k)
    print(f"BM Neumann gradient, k = {k}: GPU", flush=True)
    _, _, gg = _gpu(v, t, k, psi, pts, want_k=False)
    ops = _fp64_ops(v, t, k, pts)
    rng = np.random.default_rng(7)
    scale = np.linalg.norm(psi) / np.sqrt(len(psi))
    ad, fd = [], []
    for _ in range(6):
        d = (rng.standard_normal(len(psi)) + 1j * rng.standard_normal(len(psi))) * scale
        h = 1e-3
        fd.append((_loss64(ops, psi + h * d) - _loss64(ops, psi - h * d)) / (2 * h))
        ad.append(float(np.real(np.vdot(gg.astype(np.complex128), d))))
    ad, fd = np.array(ad), np.array(fd)
    rel = float(np.linalg.norm(ad - fd) / np.linalg.norm(fd))
    print(f"  directional derivatives {ad} vs FP64 {fd}: {rel:.2e}", flush=True)
    assert rel < 1e-4, rel


def test_vertex_and_listener_gradients_are_refused():
    from neuralat_amd.bem.grad import burton_miller_pressure

    v, t, psi, pts = _case(7.0)
    dev = torch.device("cuda:0")
    vt = torch.tensor(v, device=dev, requires_grad=True)
    with pytest.raises(NotImplementedError):
        burton_miller_pressure(vt, torch.tensor(t, device=dev), 7.0, torch.tensor(psi.astype(np.complex64),
                                                                                   device=dev),
                               torch.tensor(pts.astype(np.float32), device=dev))
    pt = torch.tensor(pts.astype(np.float32), device=dev, requires_grad=True)
    with pytest.raises(NotImplementedError):
        burton_miller_pressure(torch.tensor(v, device=dev), torch.tensor(t, device=dev), 7.0,
                               torch.tensor(psi.astype(np.complex64), device=dev), pt)
