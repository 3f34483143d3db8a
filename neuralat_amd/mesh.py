"""Synthetic surface meshes and analytic fields for the BEM workloads.

Meshes are closed, outward-oriented triangle surfaces: the normal of a
triangle (v0, v1, v2) is (v1 - v0) x (v2 - v0), the convention of the
reference's ``preprocess`` (src/bem/solver.py:40-58).

* ``icosphere(level)``  -- subdivided icosahedron, N = 20*4^L triangles,
  V = 10*4^L + 2 vertices.
* ``octasphere(level)`` -- subdivided octahedron, the family the reference's
  matrix check uses (``bempp.api.shapes.regular_sphere``,
  scripts/check_with_bempp.py:18).
* ``sh_perturb``        -- seeded radial spherical-harmonic perturbation.
* ``monopole_*``        -- the field of a point source e^{ikr}/(4 pi r), the
  analytic solution the reference's solution-level test uses
  (scripts/unit_test.py:26-39).
"""
from __future__ import annotations

import numpy as np


def _subdivide(verts: list, faces: np.ndarray, radius_project: bool = True):
    cache = {}
    vlist = list(verts)

    def mid(a, b):
        key = (a, b) if a < b else (b, a)
        if key in cache:
            return cache[key]
        m = 0.5 * (np.asarray(vlist[a]) + np.asarray(vlist[b]))
        m = m / np.linalg.norm(m)
        vlist.append(m)
        cache[key] = len(vlist) - 1
        return cache[key]

    out = []
    for a, b, c in faces:
        ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
        out += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
    return vlist, np.asarray(out, dtype=np.int64)


def _orient_outward(verts: np.ndarray, faces: np.ndarray) -> np.ndarray:
    v0, v1, v2 = verts[faces[:, 0]], verts[faces[:, 1]], verts[faces[:, 2]]
    n = np.cross(v1 - v0, v2 - v0)
    c = (v0 + v1 + v2) / 3.0
    flip = np.einsum("ij,ij->i", n, c) < 0
    faces = faces.copy()
    faces[flip, 1], faces[flip, 2] = faces[flip, 2].copy(), faces[flip, 1].copy()
    return faces


def icosphere(level: int, radius: float = 1.0):
    """Unit icosphere subdivided ``level`` times, scaled to ``radius``.
    Returns (vertices float64 [V,3], triangles int32 [N,3])."""
    t = (1.0 + 5 ** 0.5) / 2.0
    base = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t],
                     [0, 1, -t], [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    base /= np.linalg.norm(base, axis=1, keepdims=True)
    faces = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4],
                      [11, 10, 2], [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8],
                      [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    verts = list(base)
    for _ in range(level):
        verts, faces = _subdivide(verts, faces)
    v = np.asarray(verts, dtype=np.float64) * radius
    f = _orient_outward(v, faces)
    return v, f.astype(np.int32)


def octasphere(level: int, radius: float = 1.0):
    """Unit octahedron subdivided ``level`` times (projected to the sphere)."""
    base = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], dtype=np.float64)
    faces = np.array([[0, 2, 4], [2, 1, 4], [1, 3, 4], [3, 0, 4], [2, 0, 5], [1, 2, 5], [3, 1, 5], [0, 3, 5]],
                     dtype=np.int64)
    verts = list(base)
    for _ in range(level):
        verts, faces = _subdivide(verts, faces)
    v = np.asarray(verts, dtype=np.float64) * radius
    f = _orient_outward(v, faces)
    return v, f.astype(np.int32)


def sh_perturb(verts: np.ndarray, degree: int = 8, amplitude: float = 0.05, seed: int = 0) -> np.ndarray:
    """Radial perturbation r -> r (1 + s(theta, phi)) with s a seeded random
    real spherical-harmonic series of degree 1..``degree`` scaled so that
    max |s| over the vertices equals ``amplitude``."""
    from scipy.special import sph_harm_y

    rng = np.random.default_rng(seed)
    r = np.linalg.norm(verts, axis=1)
    theta = np.arccos(np.clip(verts[:, 2] / r, -1.0, 1.0))
    phi = np.arctan2(verts[:, 1], verts[:, 0])
    s = np.zeros(verts.shape[0])
    for l in range(1, degree + 1):
        for m in range(0, l + 1):
            y = sph_harm_y(l, m, theta, phi)
            s += rng.standard_normal() * y.real / l
            if m > 0:
                s += rng.standard_normal() * y.imag / l
    s *= amplitude / np.max(np.abs(s))
    return verts * (1.0 + s)[:, None]


def anisotropic_scale(verts: np.ndarray, scale) -> np.ndarray:
    return verts * np.asarray(scale, dtype=np.float64)[None, :]


# ---- analytic fields ------------------------------------------------------------

def monopole_value(x: np.ndarray, src: np.ndarray, k: float) -> np.ndarray:
    """G(x, s) = e^{ik|x-s|} / (4 pi |x-s|) for points x [...,3]."""
    d = np.linalg.norm(np.asarray(x) - np.asarray(src), axis=-1)
    return np.exp(1j * k * d) / (4 * np.pi * d)


def monopole_normal_derivative(x: np.ndarray, n: np.ndarray, src: np.ndarray, k: float) -> np.ndarray:
    """d/dn_x of G(x, s): G (ik - 1/r) (x - s).n / r."""
    d = np.asarray(x) - np.asarray(src)
    r = np.linalg.norm(d, axis=-1)
    g = np.exp(1j * k * r) / (4 * np.pi * r)
    return g * (1j * k - 1.0 / r) * np.einsum("...i,...i->...", d, n) / r


# 6-point rule used to project fields onto the discrete spaces
_A6, _B6 = 0.445948490915965, 0.091576213509771
_W6A, _W6B = 0.223381589678011, 0.109951743655322
_RULE6 = np.array([[_A6, _A6, _W6A], [_B6, _B6, _W6B], [1 - 2 * _A6, _A6, _W6A], [_A6, 1 - 2 * _A6, _W6A],
                   [1 - 2 * _B6, _B6, _W6B], [_B6, 1 - 2 * _B6, _W6B]])


def _tri_points(verts, tris, rule=_RULE6):
    v0, v1, v2 = verts[tris[:, 0]], verts[tris[:, 1]], verts[tris[:, 2]]
    x1, x2, w = rule[:, 0], rule[:, 1], rule[:, 2]
    pts = (1 - x1 - x2)[None, :, None] * v0[:, None] + x1[None, :, None] * v1[:, None] + x2[None, :, None] * v2[:, None]
    nrm = np.cross(v1 - v0, v2 - v0)
    jac = np.linalg.norm(nrm, axis=1)
    nrm = nrm / jac[:, None]
    phi = np.stack([1 - x1 - x2, x1, x2], axis=1)  # [q, 3]
    return pts, nrm, jac, w, phi


def sources_field(sources, amplitudes, k):
    """Return (value_fn, normal_derivative_fn) of a sum of monopoles."""
    sources = np.asarray(sources, dtype=np.float64).reshape(-1, 3)
    amps = np.ones(len(sources)) if amplitudes is None else np.asarray(amplitudes)

    def val(x):
        return sum(a * monopole_value(x, s, k) for s, a in zip(sources, amps))

    def dn(x, n):
        return sum(a * monopole_normal_derivative(x, n, s, k) for s, a in zip(sources, amps))

    return val, dn


def project_dp0(verts, tris, fn_of_x_n):
    """L2 projection onto piecewise constants: mean of f over each triangle
    (6-point rule)."""
    pts, nrm, jac, w, _ = _tri_points(np.asarray(verts, np.float64), np.asarray(tris))
    vals = fn_of_x_n(pts, np.broadcast_to(nrm[:, None, :], pts.shape))
    return (vals * w[None, :]).sum(axis=1)


def project_p1(verts, tris, fn_of_x_n):
    """L2 projection onto continuous piecewise linears: solve M c = (f, phi)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla

    verts = np.asarray(verts, np.float64)
    tris = np.asarray(tris)
    pts, nrm, jac, w, phi = _tri_points(verts, tris)
    vals = fn_of_x_n(pts, np.broadcast_to(nrm[:, None, :], pts.shape))
    nv = verts.shape[0]
    rhs = np.zeros(nv, np.complex128)
    for a in range(3):
        np.add.at(rhs, tris[:, a], 0.5 * jac * (vals * w[None, :] * phi[None, :, a]).sum(axis=1))
    rows, cols, data = [], [], []
    for a in range(3):
        for b in range(3):
            rows.append(tris[:, a]); cols.append(tris[:, b])
            data.append(0.5 * jac * (1.0 + (a == b)) / 12.0)
    M = sp.csr_matrix((np.concatenate(data), (np.concatenate(rows), np.concatenate(cols))), shape=(nv, nv))
    return spla.spsolve(M.tocsc(), rhs)


def random_exterior_points(n: int, rmin: float, rmax: float, seed: int = 0) -> np.ndarray:
    """n points with radius uniform in [rmin, rmax] and direction uniform on S^2."""
    rng = np.random.default_rng(seed)
    d = rng.standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = rng.uniform(rmin, rmax, size=n)
    return d * r[:, None]


def tet_box(nx: int, ny: int, nz: int, size=(1.0, 1.0, 1.0), jitter: float = 0.0, seed: int = 0):
    """Tetrahedral mesh of a box centred at the origin: nx x ny x nz cells,
    each split into 6 positively oriented tetrahedra around its main diagonal
    (conforming across cells).  ``jitter`` moves interior vertices by up to
    that fraction of a cell.  Returns (vertices [V,3] f64, tets [T,4] i32)."""
    size = np.asarray(size, dtype=np.float64)
    g = np.stack(np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij"), -1)
    verts = (g.reshape(-1, 3) / np.array([nx, ny, nz]) - 0.5) * size
    if jitter > 0:
        rng = np.random.default_rng(seed)
        inner = np.all((g.reshape(-1, 3) > 0) & (g.reshape(-1, 3) < np.array([nx, ny, nz])), axis=1)
        cell = size / np.array([nx, ny, nz])
        verts[inner] += rng.uniform(-jitter, jitter, (int(inner.sum()), 3)) * cell

    def vid(i, j, k):
        return (i * (ny + 1) + j) * (nz + 1) + k

    tets = []
    perms = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                for p in perms:
                    pos = np.array([i, j, k])
                    path = [pos.copy()]
                    for ax in p:
                        pos = pos.copy()
                        pos[ax] += 1
                        path.append(pos)
                    tets.append([vid(*q) for q in path])
    tets = np.array(tets, dtype=np.int32)
    p = verts[tets]
    det = np.einsum("ti,ti->t", p[:, 1] - p[:, 0], np.cross(p[:, 2] - p[:, 0], p[:, 3] - p[:, 0]))
    flip = det < 0
    tets[flip, 2], tets[flip, 3] = tets[flip, 3].copy(), tets[flip, 2].copy()
    return verts, tets


def tet_boundary(verts: np.ndarray, tets: np.ndarray):
    """Boundary of a tetrahedral mesh: faces that belong to one tetrahedron,
    oriented outward (away from the tetrahedron's fourth vertex).  Returns
    (surface vertices [S,3], surface triangles [F,3] i32 into them)."""
    faces = {}
    for t in np.asarray(tets):
        for skip in range(4):
            f = tuple(int(t[q]) for q in range(4) if q != skip)
            key = tuple(sorted(f))
            faces[key] = None if key in faces else (f, int(t[skip]))
    tris = []
    for val in faces.values():
        if val is None:
            continue
        (a, b, c), d = val
        n = np.cross(verts[b] - verts[a], verts[c] - verts[a])
        if np.dot(n, verts[d] - verts[a]) > 0:
            b, c = c, b
        tris.append((a, b, c))
    tris = np.array(tris, dtype=np.int64)
    used = np.unique(tris)
    remap = -np.ones(len(verts), dtype=np.int64)
    remap[used] = np.arange(len(used))
    return verts[used].copy(), remap[tris].astype(np.int32)
