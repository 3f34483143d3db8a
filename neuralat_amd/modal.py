"""FEM modal analysis: the solid's vibration modes and the per-triangle
Neumann data they give the BEM (the step upstream of the Helmholtz path).

Mirrors src/modalobj/fem.py:7-166 (``MatSet``, ``Material``, ``FEMmodel``
with ``mass_matrix`` / ``stiffness_matrix`` as torch sparse tensors,
``LOBPCG_solver``) and src/modalobj/model.py:113-160 (``VibrationObj``:
``normalize``, ``modal_analysis``, ``get_triangle_neumann``, frequencies).
The matrices are assembled by the library's HIP kernels as 3 x 3 block CSR
(``bem_fem_*``, deterministic gather, no COO duplicates); ``FEMmodel.matmul``
applies them with the library's block product.  The tetrahedralisation of a
surface mesh (FloatTetwild in the reference, src/modalobj/mesh_process.py:
21-44) and mesh-file loading stay outside: ``VibrationObj`` takes arrays.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib

SPEED_OF_SOUND = 343.2  # model.py:160


class MatSet:
    """(density, Young's modulus, Poisson's ratio, Rayleigh alpha, beta)
    per material (fem.py:23-30)."""

    Ceramic = 2700, 7.2e10, 0.19, 6, 1e-7
    Glass = 2600, 6.2e10, 0.20, 1, 1e-7
    Wood = 750, 1.1e10, 0.25, 60, 2e-6
    Plastic = 1070, 1.4e9, 0.35, 30, 1e-6
    Iron = 8000, 2.1e11, 0.28, 5, 1e-7
    Polycarbonate = 1190, 2.4e9, 0.37, 0.5, 4e-7
    Steel = 7850, 2.0e11, 0.29, 5, 3e-8
    Tin = 7265, 5e10, 0.325, 2, 3e-8

    @staticmethod
    def all():
        return [MatSet.Ceramic, MatSet.Glass, MatSet.Wood, MatSet.Plastic, MatSet.Iron, MatSet.Polycarbonate,
                MatSet.Steel, MatSet.Tin]

    @staticmethod
    def random_material(rng=None):
        lst = MatSet.all()
        idx = np.random.randint(0, len(lst)) if rng is None else int(rng.integers(0, len(lst)))
        return Material(lst[idx])


class Material:
    """Material constants (fem.py:66-73; the attribute names are the
    reference's, ``poison`` included)."""

    def __init__(self, material):
        self.density, self.youngs, self.poison, self.alpha, self.beta = material

    def relative_freq(self):
        return (self.youngs / self.density / (MatSet.Ceramic[1] / MatSet.Ceramic[0])) ** 0.5


def _check(rc, what):
    if rc != _lib.BEM_OK:
        msg = _lib.load().bem_fem_last_error().decode()
        if rc in (_lib.BEM_ERR_BAD_ARG, _lib.BEM_ERR_BAD_MESH):
            raise ValueError(f"{what}: {msg}")
        raise _lib.BemError(f"{what}: {msg} (status {rc})")


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class FEMmodel:
    """Linear-tetrahedron mass and stiffness matrices of a solid
    (fem.py:95-166).  dof 3 v + c."""

    def __init__(self, vertices, tets, material=Material(MatSet.Ceramic), device="cuda"):
        self.vertices = np.ascontiguousarray(vertices, dtype=np.float32)
        self.tets = np.ascontiguousarray(tets, dtype=np.int32)
        if self.vertices.ndim != 2 or self.vertices.shape[1] != 3 or self.tets.ndim != 2 or self.tets.shape[1] != 4:
            raise ValueError("vertices must be [V, 3] and tets [T, 4]")
        self.material = material
        self.device = torch.device(device)
        lib = _lib.load()
        h = ctypes.c_void_p()
        _check(lib.bem_fem_create(self.vertices.ctypes.data, len(self.vertices), self.tets.ctypes.data,
                                  len(self.tets), ctypes.byref(h)), "bem_fem_create")
        self._h = h
        self.nblocks = int(lib.bem_fem_num_blocks(h))
        nv = len(self.vertices)
        self.row_ptr = torch.empty(nv + 1, dtype=torch.int32, device=self.device)
        self.block_rows = torch.empty(self.nblocks, dtype=torch.int32, device=self.device)
        self.block_cols = torch.empty(self.nblocks, dtype=torch.int32, device=self.device)
        _check(lib.bem_fem_pattern(h, self.row_ptr.data_ptr(), self.block_rows.data_ptr(),
                                   self.block_cols.data_ptr(), _stream()), "bem_fem_pattern")
        self._blocks = {}
        self.stiffness_matrix_, self.mass_matrix_ = None, None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.load().bem_fem_free(h)
            except Exception:
                pass
            self._h = None

    @property
    def ndof(self):
        return 3 * len(self.vertices)

    def blocks(self, which, dtype=torch.float32):
        """Block values [nblocks, 3, 3] of "mass" or "stiffness", float32 (the
        reference's type) or float64 (both summed in FP64)."""
        if dtype not in (torch.float32, torch.float64):
            raise ValueError("dtype must be float32 or float64")
        key = (which, dtype)
        if key not in self._blocks:
            vals = torch.empty((self.nblocks, 3, 3), dtype=dtype, device=self.device)
            m = self.material
            _check(_lib.load().bem_fem_assemble(self._h, 0 if which == "mass" else 1, float(m.density),
                                                float(m.youngs), float(m.poison), int(dtype == torch.float64),
                                                vals.data_ptr(), _stream()), "bem_fem_assemble")
            self._blocks[key] = vals
        return self._blocks[key]

    def _sparse(self, which):
        r = self.block_rows.long()[:, None, None] * 3 + torch.arange(3, device=self.device)[None, :, None]
        c = self.block_cols.long()[:, None, None] * 3 + torch.arange(3, device=self.device)[None, None, :]
        r, c = torch.broadcast_tensors(r, c)
        idx = torch.stack([r.reshape(-1), c.reshape(-1)])
        return torch.sparse_coo_tensor(idx, self.blocks(which).reshape(-1), (self.ndof, self.ndof)).coalesce()

    @property
    def mass_matrix(self):
        if self.mass_matrix_ is None:
            self.mass_matrix_ = self._sparse("mass")
        return self.mass_matrix_

    @property
    def stiffness_matrix(self):
        if self.stiffness_matrix_ is None:
            self.stiffness_matrix_ = self._sparse("stiffness")
        return self.stiffness_matrix_

    def matmul(self, which, X, out=None):
        """A X for A the "mass" or "stiffness" matrix and X [3 V, m] float32 or
        float64 (the matrix in the same type)."""
        dt = torch.float64 if X.dtype == torch.float64 else torch.float32
        X = X.to(dt).contiguous()
        if X.dim() != 2 or X.shape[0] != self.ndof:
            raise ValueError(f"X must be [{self.ndof}, m]")
        if out is None:
            out = torch.empty_like(X)
        _check(_lib.load().bem_fem_spmm(self._h, int(dt == torch.float64), self.blocks(which, dt).data_ptr(),
                                        X.data_ptr(), X.stride(0), X.shape[1], out.data_ptr(), out.stride(0),
                                        _stream()), "bem_fem_spmm")
        return out


def LOBPCG_solver(stiff_matrix, mass_matrix, k):
    """The k lowest non-rigid modes (fem.py:16-20): torch.lobpcg for the
    k + 6 smallest eigenpairs of (K, M), the six rigid-body modes dropped.
    Returns (eigenvalues [k], eigenvectors [3 V, k]) as numpy."""
    vals, vecs = torch.lobpcg(stiff_matrix, 6 + k, mass_matrix, tracker=None, largest=False)
    return vals.cpu().numpy()[6:], vecs.cpu().numpy()[:, 6:]


def _eigh(G):
    """Symmetric eigendecomposition of a small Gram matrix on the host (LAPACK,
    FP64): the device solver has failed to converge on the clustered spectra
    these matrices have once block vectors converge."""
    if not bool(torch.isfinite(G).all()):
        raise FloatingPointError("non-finite Gram matrix in LOBPCG")
    w, V = torch.linalg.eigh(0.5 * (G + G.T).cpu())
    return w.to(G.device), V.to(G.device)


def _svqb(S, BS, drop=1e-10):
    """B-orthonormal basis of span(S) (S^T B S = I), dropping directions whose
    normalised Gram eigenvalue is below ``drop`` (SVQB: Stathopoulos & Wu
    2002).  Returns (S T, B S T, T)."""
    G = S.T @ BS
    G = 0.5 * (G + G.T)
    dg = torch.sqrt(torch.clamp(torch.diagonal(G), min=1e-300))
    w, V = _eigh(G / dg[:, None] / dg[None, :])
    keep = w > drop * w.max()
    T = (V[:, keep] / torch.sqrt(w[keep])[None, :]) / dg[:, None]
    return S @ T, BS @ T, T


def lobpcg(fem, k, tol=1e-6, maxiter=2000, guard=None, seed=0, jacobi=True, refresh=25):
    """The k lowest non-rigid modes of (K, M) of ``fem`` by block LOBPCG
    (Knyazev 2001) on the library's block products (FP64 products, basis and
    Rayleigh-Ritz; in FP32 the stiffness product's rounding, about 2^-24 of
    the largest eigenvalue, stalls the lowest modes' residuals near 1e-3):
    block of k + 6 + guard vectors, Jacobi
    preconditioner (the inverse diagonal of K), SVQB orthonormalisation of the
    search directions against the current block, the products of the block
    recomputed every ``refresh`` iterations.  Stops when the wanted k + 6
    residuals satisfy |K x - w M x| <= tol * w_max * |M x| (w_max: the largest
    wanted eigenvalue).  Returns (eigenvalues [k], eigenvectors [3 V, k], info)
    in FP64 numpy, the six rigid-body modes dropped."""
    f64 = torch.float64
    want = k + 6
    nb = want + (guard if guard is not None else max(4, want // 4))
    n = fem.ndof
    if n < 3 * nb:
        raise ValueError(f"the mesh has {n} degrees of freedom, too few for {k} modes")

    def A(X):
        return fem.matmul("stiffness", X)

    def B(X):
        return fem.matmul("mass", X)

    if jacobi:
        diag_blocks = fem.blocks("stiffness", f64)[fem.block_rows == fem.block_cols]  # one per vertex, in row order
        if diag_blocks.shape[0] * 3 != n:
            raise ValueError("every vertex must belong to a tetrahedron")
        dinv = (1.0 / torch.diagonal(diag_blocks, dim1=1, dim2=2).reshape(-1))[:, None]

    def rel_res(AX, BX, lam):
        R = AX[:, :want] - BX[:, :want] * lam[None, :want]
        scale = lam[want - 1].abs() * torch.linalg.vector_norm(BX[:, :want], dim=0)
        return torch.linalg.vector_norm(R, dim=0) / scale, AX - BX * lam[None, :]

    gen = torch.Generator().manual_seed(seed)
    X = torch.randn(n, nb, generator=gen, dtype=f64).to(fem.device)
    X, BX, _ = _svqb(X, B(X))
    AX = A(X)
    lam, C = _eigh(X.T @ AX)
    X, AX, BX = X @ C, AX @ C, BX @ C
    P = AP = BP = None
    it = 0
    for it in range(1, maxiter + 1):
        if it % refresh == 0:  # products of the block recomputed (drift of the recurrences)
            X, BX, T = _svqb(X, B(X))
            AX = A(X)
            lam, C = _eigh(X.T @ AX)
            X, AX, BX = X @ C, AX @ C, BX @ C
            P = AP = BP = None
        res, R = rel_res(AX, BX, lam)
        if bool((res <= tol).all()):
            break
        W = R * dinv if jacobi else R
        parts, bparts, aparts = [W], [B(W)], [A(W)]
        if P is not None:
            parts.append(P)
            bparts.append(BP)
            aparts.append(AP)
        S, BS, AS = torch.cat(parts, 1), torch.cat(bparts, 1), torch.cat(aparts, 1)
        for _ in range(2):  # B-orthogonal to the block
            cf = X.T @ BS
            S, BS, AS = S - X @ cf, BS - BX @ cf, AS - AX @ cf
        Q, BQ, T = _svqb(S, BS)
        AQ = AS @ T
        XAQ = X.T @ AQ
        G = torch.cat([torch.cat([torch.diag(lam), XAQ], 1), torch.cat([XAQ.T, Q.T @ AQ], 1)], 0)
        w, V = _eigh(G)
        CX, CQ = V[:nb, :nb], V[nb:, :nb]
        P, AP, BP = Q @ CQ, AQ @ CQ, BQ @ CQ
        X, AX, BX = X @ CX + P, AX @ CX + AP, BX @ CX + BP
        lam = w[:nb]
    res, _ = rel_res(AX, BX, lam)
    info = {"iterations": it, "max_rel_residual": float(res.max()), "converged": bool((res <= tol).all())}
    return lam[6:want].cpu().numpy(), X[:, 6:want].cpu().numpy(), info


def update_triangle_normals(vertices, triangles):
    """Unit normals (v1 - v0) x (v2 - v0) (mesh_process.py:9-19)."""
    v0, v1, v2 = vertices[triangles[:, 0]], vertices[triangles[:, 1]], vertices[triangles[:, 2]]
    n = np.cross(v1 - v0, v2 - v0)
    return n / np.linalg.norm(n, axis=1, keepdims=True)


class VibrationObj:
    """A vibrating solid (model.py:113-160): tetrahedral mesh + its surface
    mesh, given as arrays."""

    def __init__(self, vertices, tets, surf_vertices, surf_triangles):
        self.vertices = np.asarray(vertices, dtype=np.float64)
        self.tets = np.asarray(tets, dtype=np.int32)
        self.surf_vertices = np.asarray(surf_vertices, dtype=np.float64)
        self.surf_triangles = np.asarray(surf_triangles, dtype=np.int64)
        self.bbox_min = self.vertices.min(axis=0)
        self.bbox_max = self.vertices.max(axis=0)
        self.surf_normals = update_triangle_normals(self.surf_vertices, self.surf_triangles)

    def normalize(self, scale=1.0):
        """Centre the bounding box and scale its longest side to ``scale``
        (model.py:125-139)."""
        c = (self.bbox_max + self.bbox_min) / 2
        s = (self.bbox_max - self.bbox_min).max()
        self.vertices = (self.vertices - c) / s * scale
        self.surf_vertices = (self.surf_vertices - c) / s * scale
        self.bbox_min = self.vertices.min(axis=0)
        self.bbox_max = self.vertices.max(axis=0)
        self.size = (self.bbox_max - self.bbox_min).max()
        self.center = (self.bbox_max + self.bbox_min) / 2

    def modal_analysis(self, k=32, material=Material(MatSet.Plastic), solver="lobpcg"):
        """Modes of the solid and their values at the surface vertices (the
        nearest volume vertex of each, model.py:141-150).  ``solver``:
        "lobpcg" (this module's, on the library's block products, residuals
        to 1e-6) or "torch" (torch.lobpcg as LOBPCG_solver, the reference's
        choice)."""
        from scipy.spatial import KDTree

        self.fem_model = FEMmodel(self.vertices, self.tets, material)
        if solver == "torch":
            eigenvalues, eigenvectors = LOBPCG_solver(self.fem_model.stiffness_matrix, self.fem_model.mass_matrix, k)
        else:
            eigenvalues, eigenvectors, self.solve_info = lobpcg(self.fem_model, k)
        eigenvectors = eigenvectors.reshape(-1, 3, k)
        _, idx = KDTree(self.vertices).query(self.surf_vertices)
        self.modes = eigenvectors[idx].reshape(-1, 3, k)
        self.eigenvalues = eigenvalues

    def get_triangle_neumann(self, mode_id):
        """Normal velocity of mode ``mode_id`` per surface triangle: the mean
        of its vertices' mode vectors dotted with the normal (model.py:152-156)."""
        vertex_modes = self.modes[:, :, mode_id]
        return (vertex_modes[self.surf_triangles].mean(axis=1) * self.surf_normals).sum(axis=1)

    def get_frequencies(self):
        return self.eigenvalues ** 0.5 / (2 * np.pi)

    def get_frequency(self, mode_id):
        return self.eigenvalues[mode_id] ** 0.5 / (2 * np.pi)

    def get_omega(self, mode_id):
        return self.eigenvalues[mode_id] ** 0.5

    def get_wave_number(self, mode_id):
        return self.get_omega(mode_id) / SPEED_OF_SOUND


__all__ = ["MatSet", "Material", "FEMmodel", "LOBPCG_solver", "lobpcg", "VibrationObj", "update_triangle_normals"]
