"""CPU oracle for the BEM Helmholtz hot path -- TEST INFRASTRUCTURE ONLY.

This package restates the reference's algorithm on the CPU so that the HIP
product path can be checked against it.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it, and only as the checker (or, for ``cpu_baseline``, as the timed CPU
stand-in).  The product package ``neuralat_amd`` never imports it.

Contents and the reference code each part restates:

* ``assemble``   -- Galerkin boundary matrices, P1 (vertex) or DP0 (triangle)
  basis: ``src/bem/cuda/bind.cu:49-154`` (pair classification, P1 scatter),
  ``src/bem/cuda/include/integrand.h:63-394`` (regular and Sauter-Schwab
  quadrature), ``potential.h:18-65`` (kernels, r-clamp).  Implemented in C
  (``bem_oracle.c``), FP64 and an FP32 mirror.
* ``entries``    -- DP0 entries at listed (row, column) pairs, and
  ``p1_rows`` -- P1 rows of listed vertices: the same values as ``assemble``
  (pinned against it in tests/test_oracle_invariants.py).
* ``potential``  -- surface -> point operators: ``bind.cu:156-176``,
  ``integrand.h:237-255,396-406``.
* ``identity``   -- mass matrix: ``bind.cu:10-21``, ``integrand.h:257-278``.
* ``triangle2vertex`` -- ``bind.cu:23-48`` (numpy).
* ``bicgstab``   -- ``src/bem/BiCGSTAB.py:5-107`` restated (numpy).
* ``gmres``      -- textbook restarted GMRES (numpy), the solver the build adds.
* ``TRI_RULES`` / ``LINE_RULES`` -- ``src/bem/cuda/include/gauss.h:6-126``.

What pins each function (closed forms, invariants, brute force, convergence)
is listed in the docstrings of ``tests/test_oracle_*.py``.  Parity of the
reference's exact FP32 rounding and atomic summation order is unpinned (the
reference cannot be compiled or run in this image); the bar for the product
(1e-5 relative Frobenius) sits far above those effects.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_bem.so")

OPS = {"single": 0, "double": 1, "adjointdouble": 2, "hypersingular": 3, "bm_lhs": 4, "bm_rhs": 5}
BASES = {"P1": 0, "DP0": 1}

# ---- quadrature rules (gauss.h:56-126, 6-22) --------------------------------
# Triangle rules on (0,0),(1,0),(0,1), rows (xi1, xi2, w), sum(w) = 1.
_A6, _B6 = 0.445948490915965, 0.091576213509771
_W6A, _W6B = 0.223381589678011, 0.109951743655322
TRI_RULES = {
    # Strang-Fix / Cowper degree-4 rule (gauss.h:105-126 order)
    6: np.array([
        [_A6, _A6, _W6A],
        [_B6, _B6, _W6B],
        [1 - 2 * _A6, _A6, _W6A],
        [_A6, 1 - 2 * _A6, _W6A],
        [1 - 2 * _B6, _B6, _W6B],
        [_B6, 1 - 2 * _B6, _W6B],
    ]),
    3: np.array([[1 / 6, 1 / 6, 1 / 3], [2 / 3, 1 / 6, 1 / 3], [1 / 6, 2 / 3, 1 / 3]]),
    1: np.array([[1 / 3, 1 / 3, 1.0]]),
}
# Gauss-Legendre rules on [-1, 1], rows (s, w).
LINE_RULES = {n: np.stack(np.polynomial.legendre.leggauss(n), axis=1)[::-1].copy() for n in (1, 2, 3, 4)}
QUADS = {"full": (4, 6), "approx": (2, 3), "approx1": (2, 1)}  # bind.cu:180-215


def line_rule(n: int) -> np.ndarray:
    """Gauss-Legendre rule with n points on [-1, 1] as rows (s, w)."""
    return np.stack(np.polynomial.legendre.leggauss(n), axis=1)


def collapsed_tri_rule(n: int) -> np.ndarray:
    """High-order triangle rule (n*n points) by the Duffy collapse of a tensor
    Gauss-Legendre rule; rows (xi1, xi2, w) with sum(w) = 1."""
    s, w = np.polynomial.legendre.leggauss(n)
    u, wu = 0.5 * (s + 1), 0.5 * w
    U, V = np.meshgrid(u, u, indexing="ij")
    WU, WV = np.meshgrid(wu, wu, indexing="ij")
    xi1 = U
    xi2 = V * (1 - U)
    wt = 2.0 * WU * WV * (1 - U)
    return np.stack([xi1.ravel(), xi2.ravel(), wt.ravel()], axis=1)


# ---- C library ---------------------------------------------------------------
_lib = None


def build(force: bool = False) -> str:
    """Compile liboracle_bem.so with the Makefile beside this file."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE] + (["-B"] if force else []), check=True)
    return _LIB_PATH


def _load():
    global _lib
    if _lib is None:
        build()
        lib = ctypes.CDLL(_LIB_PATH)
        P = ctypes.c_void_p
        i64, i32 = ctypes.c_int64, ctypes.c_int
        for sfx, real in (("_f64", ctypes.c_double), ("_f32", ctypes.c_float)):
            f = getattr(lib, "oracle_assemble" + sfx)
            f.argtypes = [P, i64, P, i64, real, i32, i32, P, i32, P, i32, i32, i64, i64, i32, P, i32]
            f.restype = i32
            f = getattr(lib, "oracle_potential" + sfx)
            f.argtypes = [P, i64, P, i64, P, i64, real, i32, i32, P, i32, i32, P, i32]
            f.restype = i32
            f = getattr(lib, "oracle_identity" + sfx)
            f.argtypes = [P, i64, P, i64, i32, P, i32, P]
            f.restype = i32
            f = getattr(lib, "oracle_entries" + sfx)
            f.argtypes = [P, i64, P, i64, real, i32, P, i32, P, i32, i32, P, P, i64, P, i32]
            f.restype = i32
            f = getattr(lib, "oracle_p1_rows" + sfx)
            f.argtypes = [P, i64, P, i64, real, i32, P, i32, P, i32, i32, P, i64, P, i32]
            f.restype = i32
        lib.oracle_neighbors.argtypes = [P, i64, i64, P, P, P, i64, P]
        lib.oracle_neighbors.restype = ctypes.c_int64
        lib.oracle_max_threads.restype = i32
        _lib = lib
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


class OracleError(RuntimeError):
    pass


def _check(code: int):
    if code == 1:
        raise ValueError("oracle: bad argument")
    if code == 2:
        raise ValueError("oracle: bad mesh")
    if code != 0:
        raise OracleError(f"oracle error {code}")


def _prep(verts, tris, precision):
    dt = np.float64 if precision == "f64" else np.float32
    v = np.ascontiguousarray(np.asarray(verts, dtype=dt).reshape(-1, 3))
    t = np.ascontiguousarray(np.asarray(tris, dtype=np.int32).reshape(-1, 3))
    return v, t, dt


def assemble(verts, tris, k, op="single", basis="P1", quad="full", precision="f64", clamp=True,
             rows=None, parts=3, nthreads=0, tri_rule=None, line_rule_=None):
    """Dense boundary matrix as complex128.

    P1: [V, V]; DP0: [rows, N] (rows = a triangle range, default all).
    ``parts``: 1 = regular pairs only, 2 = neighbour (singular) pairs only, 3 = both.
    ``quad`` selects the reference rules (bind.cu:180-215); ``tri_rule`` /
    ``line_rule_`` override them (rows as in TRI_RULES / LINE_RULES).
    """
    lib = _load()
    v, t, dt = _prep(verts, tris, precision)
    nl, nq = QUADS[quad]
    tr = np.ascontiguousarray((TRI_RULES[nq] if tri_rule is None else tri_rule), dtype=dt)
    lr = np.ascontiguousarray((LINE_RULES[nl] if line_rule_ is None else line_rule_), dtype=dt)
    nv, nt = v.shape[0], t.shape[0]
    r0, r1 = (0, nt) if rows is None else rows
    if basis == "DP0":
        out = np.zeros((r1 - r0, nt, 2), dtype=dt)
    else:
        out = np.zeros((nv, nv, 2), dtype=dt)
    fn = getattr(lib, "oracle_assemble_" + precision)
    code = fn(_ptr(v), nv, _ptr(t), nt, float(k), OPS[op], BASES[basis], _ptr(tr), tr.shape[0], _ptr(lr),
              lr.shape[0], int(clamp), r0, r1, parts, _ptr(out), nthreads)
    _check(code)
    return (out[..., 0].astype(np.float64) + 1j * out[..., 1].astype(np.float64))


def entries(verts, tris, k, rows, cols, op="single", quad="full", precision="f64", clamp=True, nthreads=0):
    """DP0 matrix entries at the (rows[e], cols[e]) triangle pairs as
    complex128 [len(rows)]: the same values ``assemble(..., basis="DP0")``
    puts there, without building the whole matrix."""
    lib = _load()
    v, t, dt = _prep(verts, tris, precision)
    r = np.ascontiguousarray(np.asarray(rows, dtype=np.int32).reshape(-1))
    c = np.ascontiguousarray(np.asarray(cols, dtype=np.int32).reshape(-1))
    if r.shape != c.shape:
        raise ValueError("rows and cols must have the same length")
    nl, nq = QUADS[quad]
    tr = np.ascontiguousarray(TRI_RULES[nq], dtype=dt)
    lr = np.ascontiguousarray(LINE_RULES[nl], dtype=dt)
    out = np.zeros((r.shape[0], 2), dtype=dt)
    fn = getattr(lib, "oracle_entries_" + precision)
    _check(fn(_ptr(v), v.shape[0], _ptr(t), t.shape[0], float(k), OPS[op], _ptr(tr), tr.shape[0], _ptr(lr),
              lr.shape[0], int(clamp), _ptr(r), _ptr(c), r.shape[0], _ptr(out), nthreads))
    return out[:, 0].astype(np.float64) + 1j * out[:, 1].astype(np.float64)


def p1_rows(verts, tris, k, row_vertices, op="single", quad="full", precision="f64", clamp=True, nthreads=0):
    """Rows ``row_vertices`` of the P1 matrix as complex128 [len, V]: the same
    values ``assemble(..., basis="P1")`` puts in those rows, without building
    the whole matrix (cost ~ 6 N pair evaluations per row)."""
    lib = _load()
    v, t, dt = _prep(verts, tris, precision)
    rv = np.ascontiguousarray(np.asarray(row_vertices, dtype=np.int32).reshape(-1))
    nl, nq = QUADS[quad]
    tr = np.ascontiguousarray(TRI_RULES[nq], dtype=dt)
    lr = np.ascontiguousarray(LINE_RULES[nl], dtype=dt)
    out = np.zeros((rv.shape[0], v.shape[0], 2), dtype=dt)
    fn = getattr(lib, "oracle_p1_rows_" + precision)
    _check(fn(_ptr(v), v.shape[0], _ptr(t), t.shape[0], float(k), OPS[op], _ptr(tr), tr.shape[0], _ptr(lr),
              lr.shape[0], int(clamp), _ptr(rv), rv.shape[0], _ptr(out), nthreads))
    return out[..., 0].astype(np.float64) + 1j * out[..., 1].astype(np.float64)


def potential(verts, tris, points, k, op="single", basis="P1", precision="f64", clamp=True, nthreads=0,
              tri_rule=None):
    """Potential operator [P, V] (P1) or [P, N] (DP0) as complex128."""
    lib = _load()
    v, t, dt = _prep(verts, tris, precision)
    p = np.ascontiguousarray(np.asarray(points, dtype=dt).reshape(-1, 3))
    tr = np.ascontiguousarray((TRI_RULES[6] if tri_rule is None else tri_rule), dtype=dt)
    ncol = t.shape[0] if basis == "DP0" else v.shape[0]
    out = np.zeros((p.shape[0], ncol, 2), dtype=dt)
    fn = getattr(lib, "oracle_potential_" + precision)
    _check(fn(_ptr(v), v.shape[0], _ptr(t), t.shape[0], _ptr(p), p.shape[0], float(k), OPS[op], BASES[basis],
              _ptr(tr), tr.shape[0], int(clamp), _ptr(out), nthreads))
    return out[..., 0].astype(np.float64) + 1j * out[..., 1].astype(np.float64)


def identity(verts, tris, basis="P1", precision="f64", tri_rule=None):
    """Mass matrix (dense) as complex128: P1 [V, V], DP0 [N, N]."""
    lib = _load()
    v, t, dt = _prep(verts, tris, precision)
    tr = np.ascontiguousarray((TRI_RULES[6] if tri_rule is None else tri_rule), dtype=dt)
    n = t.shape[0] if basis == "DP0" else v.shape[0]
    out = np.zeros((n, n, 2), dtype=dt)
    fn = getattr(lib, "oracle_identity_" + precision)
    _check(fn(_ptr(v), v.shape[0], _ptr(t), t.shape[0], BASES[basis], _ptr(tr), tr.shape[0], _ptr(out)))
    return out[..., 0].astype(np.float64) + 1j * out[..., 1].astype(np.float64)


def neighbors(tris, nv):
    """(nb_ptr, nb_idx, nb_shared, ncolors): triangles sharing >= 1 vertex."""
    lib = _load()
    t = np.ascontiguousarray(np.asarray(tris, dtype=np.int32).reshape(-1, 3))
    nt = t.shape[0]
    cap = 64 * nt + 64
    ptr = np.zeros(nt + 1, dtype=np.int64)
    idx = np.zeros(cap, dtype=np.int32)
    sh = np.zeros(cap, dtype=np.int32)
    nc = np.zeros(1, dtype=np.int32)
    n = lib.oracle_neighbors(_ptr(t), nt, int(nv), _ptr(ptr), _ptr(idx), _ptr(sh), cap, _ptr(nc))
    if n < 0:
        raise ValueError("oracle: bad mesh")
    return ptr, idx[:n].copy(), sh[:n].copy(), int(nc[0])


def max_threads() -> int:
    return int(_load().oracle_max_threads())


# ---- numpy restatements -------------------------------------------------------

def triangle2vertex(nv, tris, features):
    """Average of triangle values over the triangles around each vertex
    (bind.cu:23-48)."""
    t = np.asarray(tris).reshape(-1, 3)
    f = np.asarray(features)
    acc = np.zeros(nv, dtype=np.complex128)
    cnt = np.zeros(nv, dtype=np.int64)
    for a in range(3):
        np.add.at(acc, t[:, a], f)
        np.add.at(cnt, t[:, a], 1)
    return acc / cnt


def bicgstab(matvec, b, x=None, nsteps=500, tol=1e-10, atol=1e-16):
    """Restatement of src/bem/BiCGSTAB.py:5-107: unpreconditioned BiCGSTAB with
    unconjugated (bilinear) rho, alpha, omega; convergence when
    ||b - A x||^2 < tol ||b||^2 or < atol, checked after every step by a fresh
    residual; stops silently when rho == 0 or nsteps run out.
    Returns (x, steps_taken, converged)."""
    b = np.asarray(b, dtype=np.complex128)
    x = np.zeros_like(b) if x is None else np.array(x, dtype=np.complex128)
    residual_tol = tol * np.vdot(b, b).real
    r = b - matvec(x)
    rr = np.vdot(r, r).real
    if rr < residual_tol or rr < atol:
        return x, 0, True
    rho_prev, alpha, omega = 1.0 + 0j, 1.0 + 0j, 1.0 + 0j
    v = np.zeros_like(b)
    p = np.zeros_like(b)
    r_hat = r.copy()
    steps = 0
    while nsteps:
        rho = np.dot(r, r_hat)
        beta = (rho / rho_prev) * (alpha / omega)
        rho_prev = rho
        p = r + beta * (p - omega * v)
        v = matvec(p)
        alpha = rho_prev / np.dot(r_hat, v)
        s = r - alpha * v
        t = matvec(s)
        omega = np.dot(t, s) / np.dot(t, t)
        x = x + alpha * p + omega * s
        steps += 1
        res = b - matvec(x)
        rr = np.vdot(res, res).real
        if rr < residual_tol or rr < atol:
            return x, steps, True
        r = s - omega * t
        if rho_prev == 0:
            break
        nsteps -= 1
    return x, steps, False


def gmres(matvec, b, tol=1e-6, restart=100, maxiter=1000, x0=None):
    """Textbook restarted GMRES(m) with modified Gram-Schmidt and Givens
    rotations, stopping on ||r|| <= tol ||b||.  Returns (x, iterations,
    converged, relative residual)."""
    b = np.asarray(b, dtype=np.complex128)
    n = b.shape[0]
    x = np.zeros(n, np.complex128) if x0 is None else np.array(x0, dtype=np.complex128)
    bnorm = np.linalg.norm(b)
    if bnorm == 0:
        return x, 0, True, 0.0
    its = 0
    while its < maxiter:
        r = b - matvec(x)
        beta = np.linalg.norm(r)
        if beta <= tol * bnorm:
            return x, its, True, beta / bnorm
        m = min(restart, maxiter - its)
        V = np.zeros((m + 1, n), np.complex128)
        H = np.zeros((m + 1, m), np.complex128)
        cs = np.zeros(m, np.complex128)
        sn = np.zeros(m, np.complex128)
        g = np.zeros(m + 1, np.complex128)
        g[0] = beta
        V[0] = r / beta
        j_done = 0
        for j in range(m):
            w = matvec(V[j])
            for i in range(j + 1):
                H[i, j] = np.vdot(V[i], w)
                w = w - H[i, j] * V[i]
            H[j + 1, j] = np.linalg.norm(w)
            if H[j + 1, j] != 0:
                V[j + 1] = w / H[j + 1, j]
            for i in range(j):
                t = cs[i] * H[i, j] + sn[i] * H[i + 1, j]
                H[i + 1, j] = -np.conj(sn[i]) * H[i, j] + np.conj(cs[i]) * H[i + 1, j]
                H[i, j] = t
            a, bb = H[j, j], H[j + 1, j]
            den = np.sqrt(abs(a) ** 2 + abs(bb) ** 2)
            # unitary rotation [[c, s], [-conj(s), conj(c)]] with c = conj(a)/den,
            # s = conj(b)/den maps (a, b) to (den, 0)
            cs[j] = np.conj(a) / den if den else 1.0
            sn[j] = np.conj(bb) / den if den else 0.0
            H[j, j] = cs[j] * a + sn[j] * bb
            H[j + 1, j] = 0
            g[j + 1] = -np.conj(sn[j]) * g[j]
            g[j] = cs[j] * g[j]
            its += 1
            j_done = j + 1
            if abs(g[j + 1]) <= tol * bnorm:
                break
        y = np.linalg.solve(np.triu(H[:j_done, :j_done]), g[:j_done])
        x = x + V[:j_done].T @ y
    r = b - matvec(x)
    rel = np.linalg.norm(r) / bnorm
    return x, its, rel <= tol * 1.0000001, rel
