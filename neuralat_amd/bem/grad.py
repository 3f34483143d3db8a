"""Differentiable exterior pressures: the DP0 CBIE solve and the P1
Burton-Miller solve.

``exterior_pressure(vertices, triangles, k, psi, points)`` returns
p = D_pot phi - S_pot psi with (M/2 - K) phi = -V psi (the reference's
CBIE + boundary2potential, src/bem/solver.py:97-104,140-147) as a
``torch.autograd.Function``: gradients flow to the vertex positions, to k and
to the Neumann data psi.  The backward pass solves the adjoint system
(M/2 - K)^H lambda = D_pot^H pbar with GMRES on the stored matrix and runs
the library's adjoint kernels, which recompute every derivative block on the
fly.  The reference has no gradients; this is checked against central finite
differences and the FP64 oracle in tests/test_gpu_grad.py.

``burton_miller_pressure(vertices, triangles, k, neumann, points)`` is the
same for the reference's production path, neumann2dirichlet +
boundary2potential in the P1 basis with the approximate rules
(src/bem/solver.py:118-126,140-147): (bm_lhs + M/2) phi = (bm_rhs -
(i/2k) M) g, p = D_pot phi - S_pot g.  Its gradients reach k and the
Neumann data g (adjoint solve with the stored matrix, the wavenumber
derivatives of the operators from bem_assemble_dk / bem_potential_dk), to
the vertices and to the listener positions (the library's P1 adjoint kernels,
which recompute every derivative block on the fly).

A forward or adjoint GMRES solve that stops at ``maxiter`` before reaching
``tol`` raises ``NotConvergedError`` (``raise_on_fail=True``, the default);
with ``raise_on_fail=False`` it warns instead.  Either way the two solves'
reports are kept in ``last_solve_info`` ({"forward": ..., "adjoint": ...}).
"""
from __future__ import annotations

import ctypes
import math
import warnings

import torch

from .. import _lib
from .solver import Mesh, _ptr, _stream, gmres, matvec

# reports of the most recent forward and adjoint solves (dicts of
# iterations, converged, rel_residual)
last_solve_info = {"forward": None, "adjoint": None}


def _report(which, info, raise_on_fail, what="exterior_pressure"):
    last_solve_info[which] = info
    if info["converged"]:
        return
    msg = (f"{what}: {which} GMRES stopped after {info['iterations']} iterations at relative "
           f"residual {info['rel_residual']:.2e}")
    if raise_on_fail:
        raise _lib.NotConvergedError(msg, info)
    warnings.warn(msg, RuntimeWarning, stacklevel=3)

_RULE6 = ((0.445948490915965, 0.445948490915965, 0.223381589678011),
          (0.091576213509771, 0.091576213509771, 0.109951743655322),
          (1 - 2 * 0.445948490915965, 0.445948490915965, 0.223381589678011),
          (0.445948490915965, 1 - 2 * 0.445948490915965, 0.223381589678011),
          (1 - 2 * 0.091576213509771, 0.091576213509771, 0.109951743655322),
          (0.091576213509771, 1 - 2 * 0.091576213509771, 0.109951743655322))


def monopole_neumann_dp0(vertices, triangles, sources, k, amplitudes=None):
    """Mean over each triangle (6-point rule) of d/dn of sum_s a_s e^{ik|x-s|}/(4 pi |x-s|),
    with torch ops so that it is differentiable in vertices, k and sources."""
    tri = triangles.long()
    v0, v1, v2 = vertices[tri[:, 0]], vertices[tri[:, 1]], vertices[tri[:, 2]]
    c = torch.cross(v1 - v0, v2 - v0, dim=1)
    n = c / torch.linalg.vector_norm(c, dim=1, keepdim=True)
    k = torch.as_tensor(k, dtype=vertices.dtype, device=vertices.device)
    out = torch.zeros(tri.shape[0], dtype=torch.complex64, device=vertices.device)
    srcs = torch.as_tensor(sources, dtype=vertices.dtype, device=vertices.device).reshape(-1, 3)
    for q, s in enumerate(srcs):
        a = 1.0 if amplitudes is None else amplitudes[q]
        for x1, x2, w in _RULE6:
            x = v0 + x1 * (v1 - v0) + x2 * (v2 - v0)
            d = x - s
            r = torch.linalg.vector_norm(d, dim=1)
            g = torch.exp(1j * k * r) / (4 * math.pi * r)
            out = out + a * w * g * (1j * k - 1.0 / r) * (d * n).sum(dim=1) / r
    return out


class _ExteriorPressure(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vertices, k, psi, triangles, points, clamp, tol, restart, maxiter, raise_on_fail, adjoint_tol):
        verts = vertices.detach().contiguous().float()
        mesh = Mesh(verts, triangles)
        kf = float(k)
        nt = mesh.nt
        dev = verts.device
        psi = psi.detach().to(torch.complex64).contiguous()
        lib = _lib.load()
        lhs = torch.empty((nt, nt), dtype=torch.complex64, device=dev)
        rhs = torch.empty(nt, dtype=torch.complex64, device=dev)
        work = torch.empty(lib.bem_cbie_work_size(mesh.handle), dtype=torch.uint8, device=dev)
        _lib.check(lib.bem_cbie_dp0(mesh.handle, kf, 0, int(clamp), _ptr(psi), _ptr(lhs), nt, _ptr(rhs),
                                    _ptr(work), _stream()), "bem_cbie_dp0")
        del work
        phi, info = gmres(lhs, rhs, tol=tol, restart=restart, maxiter=maxiter)
        last_solve_info["adjoint"] = None
        _report("forward", info, raise_on_fail)
        pts = points.detach().to(torch.float32).contiguous()
        S = _potential(mesh, pts, kf, 0, clamp)
        D = _potential(mesh, pts, kf, 1, clamp)
        p = matvec(D, phi)
        matvec(S, psi, alpha=-1.0, beta=1.0, out=p)
        ctx.mesh, ctx.lhs, ctx.S, ctx.D = mesh, lhs, S, D
        ctx.kf, ctx.clamp, ctx.tol, ctx.restart, ctx.maxiter = kf, clamp, tol, restart, maxiter
        ctx.raise_on_fail, ctx.adjoint_tol = raise_on_fail, adjoint_tol
        ctx.save_for_backward(phi, psi, pts)
        ctx.info = info
        return p

    @staticmethod
    def backward(ctx, gp):
        phi, psi, pts = ctx.saved_tensors
        mesh, lhs, S, D = ctx.mesh, ctx.lhs, ctx.S, ctx.D
        dev = phi.device
        nt = mesh.nt
        pbar = (0.5 * gp).to(torch.complex64).contiguous()
        lib = _lib.load()
        # adjoint solve (M/2 - K)^H lambda = D^H pbar, on the stored matrix
        c = matvec(D, pbar, trans=2)
        lam = torch.zeros(nt, dtype=torch.complex64, device=dev)
        restart = int(min(ctx.restart, nt))
        wk = torch.empty(lib.bem_gmres_work_size(nt, restart), dtype=torch.uint8, device=dev)
        info = _lib.SolveInfo()
        rc = lib.bem_gmres_ex(_ptr(lhs), nt, nt, 2, _ptr(c), _ptr(lam), float(ctx.adjoint_tol), restart, int(ctx.maxiter),
                              _ptr(wk), ctypes.byref(info), _stream())
        if rc not in (_lib.BEM_OK, _lib.BEM_ERR_NOT_CONVERGED):
            _lib.check(rc, "bem_gmres_ex")
        del wk
        ctx.adjoint_info = info.as_dict()
        _report("adjoint", ctx.adjoint_info, ctx.raise_on_fail)
        gvert = torch.zeros((mesh.nv, 3), dtype=torch.float64, device=dev)
        gk = torch.zeros(1, dtype=torch.float64, device=dev)
        vhl = torch.zeros((nt, 2), dtype=torch.float64, device=dev)
        _lib.check(lib.bem_adjoint_potential_dp0(mesh.handle, _ptr(pts), pts.shape[0], ctx.kf, int(ctx.clamp),
                                                 _ptr(pbar), _ptr(psi), _ptr(phi), _ptr(gvert), _ptr(gk),
                                                 _stream()), "bem_adjoint_potential_dp0")
        work = torch.empty(max(lib.bem_adjoint_work_size(mesh.handle), 1), dtype=torch.uint8, device=dev)
        _lib.check(lib.bem_adjoint_cbie_dp0(mesh.handle, ctx.kf, int(ctx.clamp), _ptr(lam), _ptr(psi), _ptr(phi),
                                            _ptr(gvert), _ptr(gk), _ptr(vhl), _ptr(work), _stream()),
                   "bem_adjoint_cbie_dp0")
        # dL/dpsi (torch convention, 2 dL/dpsi*): -2 (S^H pbar + V^H lambda)
        shp = matvec(S, pbar, trans=2)
        gpsi = -2.0 * (shp + torch.complex(vhl[:, 0], vhl[:, 1]).to(torch.complex64))
        gk_t = gk.reshape(())
        return (gvert.to(torch.float32), gk_t.to(torch.float32), gpsi, None, None, None, None, None, None, None,
                None)


def _potential(mesh, pts, k, op, clamp):
    out = torch.empty((pts.shape[0], mesh.nt), dtype=torch.complex64, device=pts.device)
    _lib.check(_lib.load().bem_potential(mesh.handle, _ptr(pts), pts.shape[0], float(k), op, 1, int(clamp),
                                         _ptr(out), mesh.nt, _stream()), "bem_potential")
    return out


def exterior_pressure(vertices, triangles, k, psi, points, clamp=True, tol=1e-6, restart=200, maxiter=2000,
                      raise_on_fail=True, adjoint_tol=None):
    """Differentiable p(points) of the DP0 CBIE solve; see the module doc.
    ``tol`` is the forward GMRES target, ``adjoint_tol`` (default: tol) the
    adjoint one: the FP32 product with A^H bottoms out near 2-4e-7 relative
    residual, so a forward target below 1e-6 wants a looser adjoint one."""
    if not torch.is_tensor(k):
        k = torch.tensor(float(k), dtype=torch.float32, device=vertices.device)
    return _ExteriorPressure.apply(vertices, k, psi, triangles, points, clamp, tol, restart, maxiter,
                                   bool(raise_on_fail), float(tol if adjoint_tol is None else adjoint_tol))


def _adjoint_solve(A, c, tol, restart, maxiter):
    """A^H x = c by GMRES on the stored matrix (bem_gmres_ex, trans = 2);
    returns (x, info dict)."""
    n = A.shape[0]
    lib = _lib.load()
    x = torch.zeros(n, dtype=torch.complex64, device=A.device)
    restart = int(min(restart, n))
    wk = torch.empty(lib.bem_gmres_work_size(n, restart), dtype=torch.uint8, device=A.device)
    info = _lib.SolveInfo()
    rc = lib.bem_gmres_ex(_ptr(A), n, A.stride(0), 2, _ptr(c), _ptr(x), float(tol), restart, int(maxiter), _ptr(wk),
                          ctypes.byref(info), _stream())
    if rc not in (_lib.BEM_OK, _lib.BEM_ERR_NOT_CONVERGED):
        _lib.check(rc, "bem_gmres_ex")
    return x, info.as_dict()


class _BurtonMillerPressure(torch.autograd.Function):
    @staticmethod
    def forward(ctx, k, neumann, vertices, triangles, points, clamp, tol, restart, maxiter, raise_on_fail,
                adjoint_tol):
        from .solver import BEM_Solver, add_identity_

        s = BEM_Solver(vertices.detach().contiguous().float(), triangles, basis="P1", clamp=clamp)
        kf = float(k)
        g = neumann.detach().to(torch.complex64).contiguous()
        L, b = s.burton_miller_system(kf, g)  # the RHS operator contracted with g, not stored
        phi, info = gmres(L, b, tol=tol, restart=restart, maxiter=maxiter)
        last_solve_info["adjoint"] = None
        _report("forward", info, raise_on_fail, "burton_miller_pressure")
        pts = points.detach().to(torch.float32).contiguous()
        S = s.potential_matrix(kf, pts, "single")
        D = s.potential_matrix(kf, pts, "double")
        p = matvec(D, phi)
        matvec(S, g, alpha=-1.0, beta=1.0, out=p)
        ctx.s, ctx.L, ctx.S, ctx.D = s, L, S, D
        ctx.kf, ctx.restart, ctx.maxiter = kf, restart, maxiter
        ctx.raise_on_fail, ctx.adjoint_tol = raise_on_fail, adjoint_tol
        ctx.save_for_backward(phi, g, pts)
        return p

    @staticmethod
    def backward(ctx, gp):
        from .solver import add_identity_

        phi, g, pts = ctx.saved_tensors
        s, L, S, D, kf = ctx.s, ctx.L, ctx.S, ctx.D, ctx.kf
        pbar = (0.5 * gp).to(torch.complex64).contiguous()  # dl/dp* (torch's gradient is 2 dl/dp*)
        # adjoint solve (bm_lhs + M/2)^H lambda = D^H pbar on the stored matrix
        lam, info = _adjoint_solve(L, matvec(D, pbar, trans=2), ctx.adjoint_tol, ctx.restart, ctx.maxiter)
        _report("adjoint", info, ctx.raise_on_fail, "burton_miller_pressure")
        grad_k = grad_g = grad_v = grad_x = None
        if ctx.needs_input_grad[2] or ctx.needs_input_grad[4]:
            # vertex (and listener) gradients from the adjoint kernels, which
            # recompute every derivative block on the fly
            lib = _lib.load()
            dev = phi.device
            gvert = torch.zeros((s.mesh.nv, 3), dtype=torch.float64, device=dev)
            gpts = torch.zeros((pts.shape[0], 3), dtype=torch.float64, device=dev) if ctx.needs_input_grad[4] else None
            if ctx.needs_input_grad[2]:
                _lib.check(lib.bem_adjoint_bm_p1(s.mesh.handle, kf, int(s.clamp), _ptr(lam), _ptr(g), _ptr(phi),
                                                 _ptr(gvert), _stream()), "bem_adjoint_bm_p1")
            _lib.check(lib.bem_adjoint_potential_p1(s.mesh.handle, _ptr(pts), pts.shape[0], kf, int(s.clamp),
                                                    _ptr(pbar), _ptr(g), _ptr(phi), _ptr(gvert),
                                                    _ptr(gpts) if gpts is not None else None, _stream()),
                       "bem_adjoint_potential_p1")
            if ctx.needs_input_grad[2]:
                grad_v = gvert.to(torch.float32)
            if gpts is not None:
                grad_x = gpts.to(torch.float32)
        if ctx.needs_input_grad[1]:
            # p = (D L^-1 R - S) g: dl/dg* = R^H lambda - S^H pbar (R assembled here:
            # the forward pass contracted it with g without storing it)
            R = add_identity_(s, s.assemble_boundary_matrix(kf, "bm_rhs", approx=True), -0.5j / kf)
            grad_g = matvec(R, lam, trans=2)
            del R
            matvec(S, pbar, trans=2, alpha=-1.0, beta=1.0, out=grad_g)
            grad_g = 2.0 * grad_g
        if ctx.needs_input_grad[0]:
            # dl/dk = 2 Re(pbar^H (D_k phi - S_k g) + lambda^H (R_k g - L_k phi)),
            # L_k = d bm_lhs/dk (the mass term does not depend on k),
            # R_k = d bm_rhs/dk + (i / 2k^2) M
            Lk = s.assemble_boundary_matrix_dk(kf, "bm_lhs", approx=True)
            Rk = add_identity_(s, s.assemble_boundary_matrix_dk(kf, "bm_rhs", approx=True), 0.5j / (kf * kf))
            t2 = matvec(Rk, g)
            matvec(Lk, phi, alpha=-1.0, beta=1.0, out=t2)
            del Lk, Rk
            Sk = s.potential_matrix_dk(kf, pts, "single")
            Dk = s.potential_matrix_dk(kf, pts, "double")
            t1 = matvec(Dk, phi)
            matvec(Sk, g, alpha=-1.0, beta=1.0, out=t1)
            gk = 2.0 * (torch.vdot(pbar, t1) + torch.vdot(lam, t2)).real
            grad_k = gk.to(torch.float32).reshape(())
        return grad_k, grad_g, grad_v, None, grad_x, None, None, None, None, None, None


def burton_miller_pressure(vertices, triangles, k, neumann, points, clamp=False, tol=1e-6, restart=200,
                           maxiter=2000, raise_on_fail=True, adjoint_tol=None):
    """Differentiable p(points) of the P1 Burton-Miller solve (the reference's
    neumann2dirichlet + boundary2potential, src/bem/solver.py:118-126,140-147):
    gradients with respect to the vertex positions, k, the Neumann data
    ``neumann`` (one complex value per vertex, held fixed when the vertices
    move) and the listener positions ``points``."""
    if not torch.is_tensor(k):
        k = torch.tensor(float(k), dtype=torch.float32, device=vertices.device)
    return _BurtonMillerPressure.apply(k, neumann, vertices, triangles, points, clamp, tol, restart, maxiter,
                                       bool(raise_on_fail), float(tol if adjoint_tol is None else adjoint_tol))
