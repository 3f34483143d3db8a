"""FP64 restatement of the reference's tetrahedral FEM matrices -- TEST
INFRASTRUCTURE ONLY (same rules as the rest of ``oracle``).

Follows src/modalobj/cuda/computeMatrix.cu literally, vectorised over the
tetrahedra with numpy:

* ``element_mass``      -- computeMatrix.cu:6-55: the 12 x 12 consistent mass
  pattern (2 on a vertex's own 3 x 3 diagonal, 1 between vertices, per
  displacement component) times density * V / 20.
* ``element_stiffness`` -- computeMatrix.cu:57-127: cofactor gradients a, b, c,
  B (6 x 12, Voigt order xx, yy, zz, xy, yz, zx) / (6 V), the isotropic D with
  the (1/2 - nu) shear terms, Ke = B^T D B V.
* ``assemble``          -- the COO triplets of every tetrahedron summed into a
  scipy CSR matrix (the reference leaves the summation to torch).
* ``modes``             -- the smallest k + 6 eigenpairs of (K, M) in FP64 by
  shift-invert Lanczos (scipy), the six rigid-body modes dropped, as
  LOBPCG_solver does (src/modalobj/fem.py:16-20).

The product (neuralat_amd/csrc/bem_fem.hip) forms the stiffness blocks from
shape-function gradients in the isotropic lambda/mu form instead; the pins in
tests/test_oracle_fem.py (rigid-body null space, total mass, the linear patch
test, symmetry, agreement of ``modes`` with a dense generalized eigensolver)
fix this restatement.
"""
from __future__ import annotations

import numpy as np


def _corners(verts, tets):
    p = np.asarray(verts, dtype=np.float64)[np.asarray(tets, dtype=np.int64)]  # [T, 4, 3]
    return p[..., 0], p[..., 1], p[..., 2]


def volumes(verts, tets):
    """|V| of each tetrahedron (computeMatrix.cu:33-34,92-93)."""
    x, y, z = _corners(verts, tets)
    V = ((x[:, 1] - x[:, 0]) * ((y[:, 2] - y[:, 0]) * (z[:, 3] - z[:, 0]) - (y[:, 3] - y[:, 0]) * (z[:, 2] - z[:, 0]))
         + (y[:, 1] - y[:, 0]) * ((x[:, 3] - x[:, 0]) * (z[:, 2] - z[:, 0]) - (x[:, 2] - x[:, 0]) * (z[:, 3] - z[:, 0]))
         + (z[:, 1] - z[:, 0]) * ((x[:, 2] - x[:, 0]) * (y[:, 3] - y[:, 0]) - (x[:, 3] - x[:, 0]) * (y[:, 2] - y[:, 0])))
    return np.abs(V) / 6.0


def element_mass(verts, tets, density):
    """[T, 12, 12] (computeMatrix.cu:35-40)."""
    V = volumes(verts, tets)
    pat = np.kron(np.ones((4, 4)) + np.eye(4), np.eye(3))
    return pat[None] * (density / 20.0 * V)[:, None, None]


def element_stiffness(verts, tets, youngs, poisson):
    """[T, 12, 12] (computeMatrix.cu:78-127)."""
    x, y, z = _corners(verts, tets)
    a = np.stack([y[:, 1] * (z[:, 3] - z[:, 2]) - y[:, 2] * (z[:, 3] - z[:, 1]) + y[:, 3] * (z[:, 2] - z[:, 1]),
                  -y[:, 0] * (z[:, 3] - z[:, 2]) + y[:, 2] * (z[:, 3] - z[:, 0]) - y[:, 3] * (z[:, 2] - z[:, 0]),
                  y[:, 0] * (z[:, 3] - z[:, 1]) - y[:, 1] * (z[:, 3] - z[:, 0]) + y[:, 3] * (z[:, 1] - z[:, 0]),
                  -y[:, 0] * (z[:, 2] - z[:, 1]) + y[:, 1] * (z[:, 2] - z[:, 0]) - y[:, 2] * (z[:, 1] - z[:, 0])], 1)
    b = np.stack([-x[:, 1] * (z[:, 3] - z[:, 2]) + x[:, 2] * (z[:, 3] - z[:, 1]) - x[:, 3] * (z[:, 2] - z[:, 1]),
                  x[:, 0] * (z[:, 3] - z[:, 2]) - x[:, 2] * (z[:, 3] - z[:, 0]) + x[:, 3] * (z[:, 2] - z[:, 0]),
                  -x[:, 0] * (z[:, 3] - z[:, 1]) + x[:, 1] * (z[:, 3] - z[:, 0]) - x[:, 3] * (z[:, 1] - z[:, 0]),
                  x[:, 0] * (z[:, 2] - z[:, 1]) - x[:, 1] * (z[:, 2] - z[:, 0]) + x[:, 2] * (z[:, 1] - z[:, 0])], 1)
    c = np.stack([x[:, 1] * (y[:, 3] - y[:, 2]) - x[:, 2] * (y[:, 3] - y[:, 1]) + x[:, 3] * (y[:, 2] - y[:, 1]),
                  -x[:, 0] * (y[:, 3] - y[:, 2]) + x[:, 2] * (y[:, 3] - y[:, 0]) - x[:, 3] * (y[:, 2] - y[:, 0]),
                  x[:, 0] * (y[:, 3] - y[:, 1]) - x[:, 1] * (y[:, 3] - y[:, 0]) + x[:, 3] * (y[:, 1] - y[:, 0]),
                  -x[:, 0] * (y[:, 2] - y[:, 1]) + x[:, 1] * (y[:, 2] - y[:, 0]) - x[:, 2] * (y[:, 1] - y[:, 0])], 1)
    V = volumes(verts, tets)
    T = len(V)
    B = np.zeros((T, 6, 12))
    for i in range(4):
        B[:, 0, 3 * i] = a[:, i]
        B[:, 1, 3 * i + 1] = b[:, i]
        B[:, 2, 3 * i + 2] = c[:, i]
        B[:, 3, 3 * i] = b[:, i]
        B[:, 3, 3 * i + 1] = a[:, i]
        B[:, 4, 3 * i + 1] = c[:, i]
        B[:, 4, 3 * i + 2] = b[:, i]
        B[:, 5, 3 * i] = c[:, i]
        B[:, 5, 3 * i + 2] = a[:, i]
    B /= (6.0 * V)[:, None, None]
    nu = float(poisson)
    D = np.zeros((6, 6))
    D[:3, :3] = nu
    D[0, 0] = D[1, 1] = D[2, 2] = 1.0 - nu
    D[3, 3] = D[4, 4] = D[5, 5] = 0.5 - nu
    D *= float(youngs) / (1.0 + nu) / (1.0 - 2.0 * nu)
    return np.einsum("tki,kl,tlj->tij", B, D, B) * V[:, None, None]


def assemble(verts, tets, kind, density=1.0, youngs=1.0, poisson=0.3):
    """Global [3 nv, 3 nv] scipy CSR matrix, kind "mass" or "stiffness"."""
    import scipy.sparse as sp

    tets = np.asarray(tets, dtype=np.int64)
    nv = len(verts)
    Ke = element_mass(verts, tets, density) if kind == "mass" else element_stiffness(verts, tets, youngs, poisson)
    dof = (3 * tets[:, :, None] + np.arange(3)[None, None, :]).reshape(-1, 12)
    rows = np.repeat(dof[:, :, None], 12, axis=2)
    cols = np.repeat(dof[:, None, :], 12, axis=1)
    return sp.coo_matrix((Ke.ravel(), (rows.ravel(), cols.ravel())), shape=(3 * nv, 3 * nv)).tocsr()


def modes(K, M, k, shift=None):
    """The k smallest non-rigid eigenpairs (K - w M) x = 0 in FP64: the k + 6
    smallest by shift-invert Lanczos, the six rigid-body modes dropped."""
    import scipy.sparse.linalg as sla

    if shift is None:
        shift = -1e-3 * abs(K.diagonal()).max() / abs(M.diagonal()).max()
    vals, vecs = sla.eigsh(K, k + 6, M, sigma=shift, which="LM")
    order = np.argsort(vals)
    return vals[order][6:], vecs[:, order][:, 6:]
