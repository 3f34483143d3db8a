"""Pins of the FEM oracle (oracle/fem.py) by what the mathematics fixes.

Any nodal displacement of a linear tetrahedron is a linear field
u(x) = G x + t (12 dofs = 9 + 3), so energy identities over random linear
fields determine each symmetric 12 x 12 element matrix completely:
* stiffness: u^T Ke u = V (lambda tr(eps)^2 + 2 mu eps:eps), eps = sym(G)
  (exact for linear elements: the patch test);
* mass: u^T Me u = rho int_T |G x + t|^2 dV, with the tetrahedron's moments
  int x x^T dV = V/20 (sum p_i p_i^T + (sum p_i)(sum p_i)^T), int x dV = V c.
On assembled matrices: the six rigid-body modes are in the null space of K,
the total mass is rho times the volume, and the FP64 eigensolver agrees with a
dense generalized eigensolver on a small mesh (six zero eigenvalues first).
"""
import numpy as np
import pytest
import scipy.linalg as sla

from neuralat_amd import mesh
from oracle import fem


def _one_tet(seed):
    rng = np.random.default_rng(seed)
    p = rng.standard_normal((4, 3))
    if np.linalg.det(np.stack([p[1] - p[0], p[2] - p[0], p[3] - p[0]])) < 0:
        p[[2, 3]] = p[[3, 2]]
    return p, np.array([[0, 1, 2, 3]])


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_element_stiffness_energy_of_linear_fields(seed):
    p, t = _one_tet(seed)
    E, nu = 7.2e10, 0.19
    lam, mu = E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))
    Ke = fem.element_stiffness(p, t, E, nu)[0]
    V = fem.volumes(p, t)[0]
    rng = np.random.default_rng(10 + seed)
    assert np.allclose(Ke, Ke.T, rtol=0, atol=1e-9 * abs(Ke).max())
    for _ in range(100):
        G, tr = rng.standard_normal((3, 3)), rng.standard_normal(3)
        u = (p @ G.T + tr).ravel()
        eps = 0.5 * (G + G.T)
        want = V * (lam * np.trace(eps) ** 2 + 2 * mu * np.sum(eps * eps))
        assert abs(u @ Ke @ u - want) <= 1e-10 * abs(want)


@pytest.mark.parametrize("seed", [0, 1])
def test_element_mass_integral_of_linear_fields(seed):
    p, t = _one_tet(seed)
    rho = 2700.0
    Me = fem.element_mass(p, t, rho)[0]
    V = fem.volumes(p, t)[0]
    S = V / 20.0 * (p.T @ p + np.outer(p.sum(0), p.sum(0)))
    m1 = V * p.mean(0)
    rng = np.random.default_rng(20 + seed)
    for _ in range(100):
        G, tr = rng.standard_normal((3, 3)), rng.standard_normal(3)
        u = (p @ G.T + tr).ravel()
        want = rho * (np.trace(G @ S @ G.T) + 2 * tr @ G @ m1 + V * tr @ tr)
        assert abs(u @ Me @ u - want) <= 1e-10 * abs(want)


def test_assembled_rigid_modes_mass_and_symmetry():
    v, t = mesh.tet_box(4, 3, 5, (0.3, 0.2, 0.5), jitter=0.25, seed=3)
    K = fem.assemble(v, t, "stiffness", youngs=2e11, poisson=0.29)
    M = fem.assemble(v, t, "mass", density=7850.0)
    assert abs(K - K.T).max() <= 1e-12 * abs(K).max()
    assert abs(M - M.T).max() <= 1e-12 * abs(M).max()
    n = len(v)
    rigid = []
    for a in range(3):
        e = np.zeros((n, 3))
        e[:, a] = 1.0
        rigid.append(e.ravel())
    for w in np.eye(3):
        rigid.append(np.cross(w, v).ravel())
    for r in rigid:
        assert np.linalg.norm(K @ r) <= 1e-9 * abs(K).max() * np.linalg.norm(r)
    vol = fem.volumes(v, t).sum()
    assert abs(rigid[0] @ (M @ rigid[0]) - 7850.0 * vol) <= 1e-9 * 7850.0 * vol


def test_modes_match_dense_generalized_eigensolver():
    v, t = mesh.tet_box(3, 2, 2, (0.3, 0.2, 0.2), jitter=0.2, seed=5)
    K = fem.assemble(v, t, "stiffness", youngs=1.4e9, poisson=0.35)
    M = fem.assemble(v, t, "mass", density=1070.0)
    w_all = sla.eigh(K.toarray(), M.toarray(), eigvals_only=True)
    assert np.all(np.abs(w_all[:6]) <= 1e-6 * w_all[6])  # six rigid-body modes
    w, x = fem.modes(K, M, 8)
    assert np.allclose(w, w_all[6:14], rtol=1e-8)
    assert np.allclose((x.T @ (M @ x)), np.eye(8), atol=1e-8)
