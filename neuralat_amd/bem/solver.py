"""Python surface of the BEM Helmholtz path, a drop-in for the reference's
``src/bem/solver.py`` (class ``BEM_Solver``, ``solve_linear_equation``,
``map_triangle2vertex``, ``check_tensor``, ``preprocess``).

Every numerical step runs in libbem_hip.so (hand-written HIP for gfx950)
through its C ABI; torch only allocates device memory and supplies the
current stream.  Differences from the reference, all opt-in keywords:

* ``basis="P1"`` (default, the reference's vertex basis) or ``"DP0"``
  (one unknown per triangle);
* ``clamp=True`` (default) reproduces the reference's r >= 1e-3 kernel clamp
  (potential.h:16-23); ``clamp=False`` evaluates the exact kernel;
* ``CBIE/HBIE/neumann2dirichlet`` solve with device GMRES (tol 1e-6 relative
  residual) unless ``solver="bicgstab"`` asks for the reference's BiCGSTAB.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from .. import _lib


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t: torch.Tensor):
    return ctypes.c_void_p(t.data_ptr())


def check_tensor(tensor, dtype):
    """Same checks as the reference (solver.py:34-37): dtype, device, layout."""
    assert tensor.dtype == dtype
    assert tensor.is_cuda
    assert tensor.is_contiguous()


class Mesh:
    """Device-resident preprocessed mesh (a ``bem_mesh`` handle)."""

    def __init__(self, vertices: torch.Tensor, triangles: torch.Tensor):
        check_tensor(vertices, torch.float32)
        check_tensor(triangles, torch.int32)
        if vertices.dim() != 2 or vertices.shape[1] != 3 or triangles.dim() != 2 or triangles.shape[1] != 3:
            raise ValueError("vertices must be [V,3] and triangles [N,3]")
        lib = _lib.load()
        h = ctypes.c_void_p()
        with torch.cuda.device(vertices.device):
            rc = lib.bem_mesh_create(_ptr(vertices), vertices.shape[0], _ptr(triangles), triangles.shape[0], 1,
                                     _stream(), ctypes.byref(h))
        _lib.check(rc, "bem_mesh_create")
        self._h = h
        self._lib = lib
        self.device = vertices.device
        self.nv = int(vertices.shape[0])
        self.nt = int(triangles.shape[0])
        npairs, maxnb = ctypes.c_int64(), ctypes.c_int32()
        _lib.check(lib.bem_mesh_info(h, None, None, ctypes.byref(npairs), ctypes.byref(maxnb)))
        self.n_neighbor_pairs = int(npairs.value)
        self.max_neighbors = int(maxnb.value)

    @property
    def handle(self):
        return self._h

    def set_vertices(self, vertices: torch.Tensor):
        """New positions [V,3] for the same triangles (bem_mesh_set_vertices)."""
        check_tensor(vertices, torch.float32)
        if tuple(vertices.shape) != (self.nv, 3):
            raise ValueError(f"vertices must be [{self.nv},3]")
        vertices = vertices.contiguous()
        _lib.check(self._lib.bem_mesh_set_vertices(self._h, _ptr(vertices), 1, _stream()), "bem_mesh_set_vertices")

    def normals_curls(self):
        n = torch.empty((self.nt, 3), dtype=torch.float32, device=self.device)
        c = torch.empty((self.nt, 3, 3), dtype=torch.float32, device=self.device)
        _lib.check(self._lib.bem_mesh_normals_curls(self._h, _ptr(n), _ptr(c), _stream()))
        return n, c

    def size(self, basis):
        return self.nt if basis == "DP0" else self.nv

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                torch.cuda.current_stream(self.device).synchronize()
            except Exception:
                pass
            self._lib.bem_mesh_destroy(h)
            self._h = None


class GeometrySet:
    """Several placements ("samples") of one mesh's triangles (a
    ``bem_geomset``): each sample's position-dependent data on the device, the
    mesh's connectivity (tiling, neighbour pairs) shared.  ``vertices``:
    [ns, V, 3].  Operators of all samples come from one launch per kernel
    (the reference builds one BEM_Solver per sample, scene.py:218-291)."""

    def __init__(self, mesh: Mesh, vertices):
        if torch.is_tensor(vertices):
            vertices = vertices.detach().cpu().numpy()
        v = np.ascontiguousarray(np.asarray(vertices, dtype=np.float32).reshape(-1, mesh.nv, 3))
        lib = _lib.load()
        h = ctypes.c_void_p()
        with torch.cuda.device(mesh.device):
            rc = lib.bem_geomset_create(mesh.handle, v.ctypes.data, v.shape[0], _stream(), ctypes.byref(h))
        _lib.check(rc, "bem_geomset_create")
        self._h, self._lib, self.mesh, self.ns = h, lib, mesh, int(v.shape[0])
        self.device = mesh.device

    @property
    def handle(self):
        return self._h

    def _ks(self, ks):
        ks = np.ascontiguousarray(np.asarray(ks, dtype=np.float32).reshape(-1))
        if ks.shape[0] != self.ns:
            raise ValueError(f"need one wavenumber per sample ({self.ns})")
        return ks

    def assemble(self, ks, layer_type, approx=False, basis="P1", clamp=True):
        """Operator of every sample at its wavenumber: complex64 [ns, n, n]."""
        if layer_type not in _lib.OPS:
            raise ValueError(f"unknown layer type {layer_type!r}")
        quad = "approx_1" if approx == "approx_1" else ("approx" if approx else "full")
        ks = self._ks(ks)
        n = self.mesh.size(basis)
        out = torch.empty((self.ns, n, n), dtype=torch.complex64, device=self.device)
        _lib.check(self._lib.bem_assemble_samples(self._h, ks.ctypes.data, _lib.OPS[layer_type], _lib.BASES[basis],
                                                  _lib.QUADS[quad], int(clamp), _ptr(out), n, n * n, _stream()),
                   "bem_assemble_samples")
        return out

    def add_identity_(self, A, alphas, basis="P1"):
        """A[z] += alphas[z] M_z in place."""
        al = np.ascontiguousarray(np.asarray(alphas, dtype=np.complex64).reshape(-1))
        if al.shape[0] != self.ns:
            raise ValueError(f"need one coefficient per sample ({self.ns})")
        al = al.view(np.float32)
        n = A.shape[-1]
        _lib.check(self._lib.bem_add_identity_samples(self._h, _lib.BASES[basis], al.ctypes.data, _ptr(A), A.stride(1),
                                                      A.stride(0), _stream()), "bem_add_identity_samples")
        return A

    def potential(self, ks, points, layer_type, basis="P1", clamp=True):
        """Potential operator of every sample at its own points [ns, np, 3]:
        complex64 [ns, np, n]."""
        ks = self._ks(ks)
        pts = torch.as_tensor(points, dtype=torch.float32, device=self.device).reshape(self.ns, -1, 3).contiguous()
        npts, n = pts.shape[1], self.mesh.size(basis)
        out = torch.empty((self.ns, npts, n), dtype=torch.complex64, device=self.device)
        _lib.check(self._lib.bem_potential_samples(self._h, _ptr(pts), npts, ks.ctypes.data, _lib.OPS[layer_type],
                                                   _lib.BASES[basis], int(clamp), _ptr(out), n, npts * n, _stream()),
                   "bem_potential_samples")
        return out

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                torch.cuda.current_stream(self.device).synchronize()
            except Exception:
                pass
            self._lib.bem_geomset_destroy(h)
            self._h = None


def matvec_strided(A: torch.Tensor, x: torch.Tensor, alpha=1.0, beta=0.0, out=None):
    """out[z] = alpha A[z] x[z] + beta out[z] for A [nz, rows, cols], x [nz, cols]
    (one launch)."""
    check_tensor(A, torch.complex64)
    check_tensor(x, torch.complex64)
    nz, rows, cols = A.shape
    if out is None:
        out = torch.empty((nz, rows), dtype=torch.complex64, device=A.device)
    a, b = complex(alpha), complex(beta)
    _lib.check(_lib.load().bem_matvec_strided(_ptr(A), rows, cols, A.stride(1), A.stride(0), _ptr(x), x.stride(0),
                                              _ptr(out), out.stride(0), nz, a.real, a.imag, b.real, b.imag,
                                              _stream()), "bem_matvec_strided")
    return out


def burton_miller_samples(gs: GeometrySet, ks, neumanns, clamp=True):
    """The neumann2dirichlet systems (solver.py:120-125, approximate rules) of
    every sample of a geometry set at its wavenumber: lhs [ns, V, V] =
    bm_lhs + M/2, rhs [ns, V] = (bm_rhs - (i/k) M/2) psi; one launch per
    kernel for all samples (bem_burton_miller_samples: the right-hand side's
    operator contracted with psi inside the LHS kernels, never stored)."""
    kh = gs._ks(ks)
    psi = torch.as_tensor(neumanns, device=gs.device).to(torch.complex64).reshape(gs.ns, -1).contiguous()
    n = psi.shape[1]
    lhs = torch.empty((gs.ns, n, n), dtype=torch.complex64, device=gs.device)
    rhs = torch.empty((gs.ns, n), dtype=torch.complex64, device=gs.device)
    _lib.check(_lib.load().bem_burton_miller_samples(gs.handle, kh.ctypes.data, _lib.QUADS["approx"], int(clamp),
                                                     _ptr(psi), _ptr(lhs), n, n * n, _ptr(rhs), _stream()),
               "bem_burton_miller_samples")
    return lhs, rhs


def boundary2potential_samples(gs: GeometrySet, ks, neumanns, dirichlets, points, clamp=True):
    """p[z] = D_z phi_z - S_z psi_z at every sample's own points [ns, np, 3]
    (solver.py:140-147): [ns, np]."""
    psi = torch.as_tensor(neumanns, device=gs.device).to(torch.complex64).reshape(gs.ns, -1).contiguous()
    phi = torch.as_tensor(dirichlets, device=gs.device).to(torch.complex64).reshape(gs.ns, -1).contiguous()
    S = gs.potential(ks, points, "single", clamp=clamp)
    D = gs.potential(ks, points, "double", clamp=clamp)
    p = matvec_strided(D, phi)
    return matvec_strided(S, psi, alpha=-1.0, beta=1.0, out=p)


def preprocess(vertices, triangles):
    """Unit normals [N,3] and P1 surface curls [N,3,3] (row a = n x grad phi_a),
    as the reference's preprocess (solver.py:40-58), computed by the library."""
    return Mesh(vertices, triangles).normals_curls()


# ---------------------------------------------------------------------------
# dense helpers on the C ABI
# ---------------------------------------------------------------------------

def matvec(A: torch.Tensor, x: torch.Tensor, trans: int = 0, alpha=1.0, beta=0.0, out=None):
    """out = alpha op(A) x + beta out with op = A (0), A^T (1), A^H (2)."""
    check_tensor(A, torch.complex64)
    check_tensor(x, torch.complex64)
    rows, cols = A.shape
    n_out = cols if trans else rows
    if out is None:
        out = torch.empty(n_out, dtype=torch.complex64, device=A.device)
    lib = _lib.load()
    wsz = lib.bem_matvec_work_size(rows, cols, trans)
    work = torch.empty(max(wsz, 1), dtype=torch.uint8, device=A.device)
    a, b = complex(alpha), complex(beta)
    _lib.check(lib.bem_matvec(_ptr(A), rows, cols, A.stride(0), trans, _ptr(x), _ptr(out), a.real, a.imag, b.real,
                              b.imag, _ptr(work), _stream()), "bem_matvec")
    return out


def combine(A, a, B, b):
    """a A + b B (A or B may be None) as a new complex64 matrix."""
    ref = A if A is not None else B
    rows, cols = ref.shape
    C = torch.empty((rows, cols), dtype=torch.complex64, device=ref.device)
    a, b = complex(a), complex(b)
    lib = _lib.load()
    _lib.check(lib.bem_combine(_ptr(A) if A is not None else None, A.stride(0) if A is not None else cols,
                               _ptr(B) if B is not None else None, B.stride(0) if B is not None else cols, _ptr(C),
                               cols, rows, cols, a.real, a.imag, b.real, b.imag, _stream()), "bem_combine")
    return C


def scale_(A: torch.Tensor, a):
    """A *= a in place."""
    a = complex(a)
    rows, cols = A.shape
    _lib.check(_lib.load().bem_combine(_ptr(A), A.stride(0), None, cols, _ptr(A), A.stride(0), rows, cols, a.real,
                                       a.imag, 0.0, 0.0, _stream()), "bem_combine")
    return A


def add_identity_(solver, A: torch.Tensor, alpha, basis=None):
    """A += alpha M in place (M the mass matrix of the solver's basis)."""
    basis = basis or solver.basis
    alpha = complex(alpha)
    _lib.check(_lib.load().bem_add_identity(solver.mesh.handle, _lib.BASES[basis], alpha.real, alpha.imag, _ptr(A),
                                            A.stride(0), _stream()), "bem_add_identity")
    return A


def gmres(A: torch.Tensor, b: torch.Tensor, x0=None, tol=1e-6, restart=100, maxiter=1000, raise_on_fail=False,
          time_products=False):
    """Device GMRES(restart) on a dense complex64 matrix.  Returns (x, info);
    with ``time_products`` info also holds the summed HIP-event time of the
    matrix-vector products ("gemv_ms") and their number ("gemv_count")."""
    check_tensor(A, torch.complex64)
    check_tensor(b, torch.complex64)
    n = A.shape[0]
    x = torch.zeros(n, dtype=torch.complex64, device=A.device) if x0 is None else x0.clone()
    lib = _lib.load()
    restart = int(min(restart, max(n, 1)))
    work = torch.empty(lib.bem_gmres_work_size(n, restart), dtype=torch.uint8, device=A.device)
    info = _lib.SolveInfo()
    ms, cnt = ctypes.c_double(), ctypes.c_int64()
    if time_products:
        rc = lib.bem_gmres_timed(_ptr(A), n, A.stride(0), 0, _ptr(b), _ptr(x), float(tol), restart, int(maxiter),
                                 _ptr(work), ctypes.byref(info), ctypes.byref(ms), ctypes.byref(cnt), _stream())
    else:
        rc = lib.bem_gmres(_ptr(A), n, A.stride(0), _ptr(b), _ptr(x), float(tol), restart, int(maxiter), _ptr(work),
                           ctypes.byref(info), _stream())
    if not (rc == _lib.BEM_ERR_NOT_CONVERGED and not raise_on_fail):
        _lib.check(rc, "bem_gmres")
    out = info.as_dict()
    if time_products:
        out["gemv_ms"] = float(ms.value)
        out["gemv_count"] = int(cnt.value)
    return x, out


def bicgstab(A: torch.Tensor, b: torch.Tensor, x=None, nsteps=500, tol=1e-10, atol=1e-16):
    """Reference BiCGSTAB semantics (BiCGSTAB.py) on a dense matrix, on the
    device.  Returns (x, info)."""
    check_tensor(A, torch.complex64)
    check_tensor(b, torch.complex64)
    n = A.shape[0]
    x = torch.zeros(n, dtype=torch.complex64, device=A.device) if x is None else x.clone()
    lib = _lib.load()
    work = torch.empty(lib.bem_bicgstab_work_size(n), dtype=torch.uint8, device=A.device)
    info = _lib.SolveInfo()
    _lib.check(lib.bem_bicgstab(_ptr(A), n, A.stride(0), _ptr(b), _ptr(x), int(nsteps), float(tol), float(atol),
                                _ptr(work), ctypes.byref(info), _stream()), "bem_bicgstab")
    return x, info.as_dict()


def _bicgstab_over(apply, b, x=None, nsteps=None, tol=1e-10, atol=1e-16):
    """BiCGSTAB (van der Vorst, SIAM J. Sci. Stat. Comput. 13, 1992, Alg. 3)
    over a caller-supplied operator ``apply(v) -> A v``, with the stopping
    behaviour of the reference's src/bem/BiCGSTAB.py:30-107:

    * the inner products of the recurrence are bilinear (no conjugation), the
      shadow residual is the initial residual;
    * after each step the TRUE residual b - A x is formed and the solve stops
      when ||b - A x||^2 < tol ||b||^2 or < atol (so one extra product per step);
    * the recurrence breaks off when its rho vanishes, and running out of steps
      returns the current iterate without a signal.

    ``nsteps`` None means len(b) steps."""
    b = b.detach().clone()
    x = torch.zeros_like(b) if x is None else x
    limit = tol * torch.vdot(b, b).real

    def done(r):
        rr = torch.vdot(r, r).real
        return bool(rr < limit or rr < atol)

    r = b - apply(x)
    if done(r):
        return x
    shadow = r.clone()
    # state of the two-term recurrence: previous rho, alpha, omega and the
    # search direction p with its image v = A p
    state = {"rho": 1.0, "alpha": 1.0, "omega": 1.0}
    p = torch.zeros_like(b)
    v = torch.zeros_like(b)
    budget = b.shape[0] if nsteps is None else nsteps
    while budget:
        rho = (r * shadow).sum()
        p = r + (rho / state["rho"]) * (state["alpha"] / state["omega"]) * (p - state["omega"] * v)
        v = apply(p)
        alpha = rho / (shadow * v).sum()
        s = r - alpha * v
        t = apply(s)
        omega = (t * s).sum() / (t * t).sum()
        x = x + alpha * p + omega * s
        if done(b - apply(x)):
            return x
        r = s - omega * t
        state.update(rho=rho, alpha=alpha, omega=omega)
        if rho == 0:
            break
        budget -= 1
    return x


def solve_linear_equation(A_func, b, x=None, nsteps=500, tol=1e-10, atol=1e-16):
    """Reference signature and semantics (solver.py:61-66): BiCGSTAB with
    ||r||^2 < tol ||b||^2 (or < atol), returning x silently on failure.  A
    dense complex64 matrix runs on the library's device BiCGSTAB; a callable
    runs the same iteration over the caller's matvec (``_bicgstab_over``)."""
    if callable(A_func):
        return _bicgstab_over(A_func, b, x=x, nsteps=nsteps, tol=tol, atol=atol)
    return bicgstab(A_func, b, x=x, nsteps=nsteps, tol=tol, atol=atol)[0]


# ---------------------------------------------------------------------------
# the solver object
# ---------------------------------------------------------------------------

class BEM_Solver:
    """Drop-in for the reference's ``BEM_Solver`` (solver.py:69-147)."""

    def __init__(self, vertices, triangles, basis="P1", clamp=True):
        if isinstance(vertices, np.ndarray):
            vertices = torch.from_numpy(vertices).cuda().float()
        if isinstance(triangles, np.ndarray):
            triangles = torch.from_numpy(triangles).cuda().int()
        check_tensor(vertices, torch.float32)
        check_tensor(triangles, torch.int32)
        if basis not in _lib.BASES:
            raise ValueError("basis must be 'P1' or 'DP0'")
        self.vertices = vertices
        self.triangles = triangles
        self.device = vertices.device
        self.basis = basis
        self.clamp = bool(clamp)
        self.mesh = Mesh(vertices, triangles)
        self.normals, self.surface_curls_trans = self.mesh.normals_curls()
        self.last_solve_info = None

    def set_vertices(self, vertices):
        """Move the surface (same triangles) without rebuilding the mesh
        handle: the next operators use the new positions."""
        if isinstance(vertices, np.ndarray):
            vertices = torch.from_numpy(vertices).to(self.device).float()
        self.mesh.set_vertices(vertices)
        self.vertices = vertices
        self.normals, self.surface_curls_trans = self.mesh.normals_curls()

    # -- operators ------------------------------------------------------------
    def _n(self, basis):
        return self.mesh.size(basis)

    def assemble_boundary_matrix(self, wavenumber, layer_type, approx=False, basis=None, clamp=None):
        """Dense operator matrix, complex64 [n, n] (bind.cu:49-154).
        ``approx``: False = 4/6-point rules, True = 2/3, "approx_1" = 2/1."""
        basis = basis or self.basis
        clamp = self.clamp if clamp is None else clamp
        if layer_type not in _lib.OPS:
            raise ValueError(f"unknown layer type {layer_type!r}")
        quad = "approx_1" if approx == "approx_1" else ("approx" if approx else "full")
        n = self._n(basis)
        out = torch.empty((n, n), dtype=torch.complex64, device=self.device)
        lib = _lib.load()
        _lib.check(lib.bem_assemble(self.mesh.handle, float(wavenumber), _lib.OPS[layer_type], _lib.BASES[basis],
                                    _lib.QUADS[quad], int(clamp), _ptr(out), n, _stream()), "bem_assemble")
        return out

    def assemble_boundary_matrix_dk(self, wavenumber, layer_type, approx=False, clamp=None):
        """d/dk of the P1 operator matrix (single, double, adjointdouble,
        bm_lhs, bm_rhs), complex64 [nv, nv]; no identity terms.  The reference
        has no derivatives: this serves the adjoint of neumann2dirichlet."""
        if self.basis != "P1":
            raise ValueError("wavenumber derivatives of the operators are built for the P1 basis")
        clamp = self.clamp if clamp is None else clamp
        if layer_type not in _lib.OPS:
            raise ValueError(f"unknown layer type {layer_type!r}")
        quad = "approx_1" if approx == "approx_1" else ("approx" if approx else "full")
        n = self._n("P1")
        out = torch.empty((n, n), dtype=torch.complex64, device=self.device)
        _lib.check(_lib.load().bem_assemble_dk(self.mesh.handle, float(wavenumber), _lib.OPS[layer_type],
                                               _lib.QUADS[quad], int(clamp), _ptr(out), n, _stream()),
                   "bem_assemble_dk")
        return out

    def assemble_boundary_matrix_batch(self, wavenumbers, layer_type, approx=False, basis=None, clamp=None):
        """The operator at every wavenumber of ``wavenumbers`` (a sequence of
        floats), complex64 [len, n, n]: one launch per kernel for up to 64 of
        them (the reference assembles one wavenumber per call)."""
        basis = basis or self.basis
        clamp = self.clamp if clamp is None else clamp
        if layer_type not in _lib.OPS:
            raise ValueError(f"unknown layer type {layer_type!r}")
        quad = "approx_1" if approx == "approx_1" else ("approx" if approx else "full")
        ks = np.ascontiguousarray(np.asarray(wavenumbers, dtype=np.float32).reshape(-1))
        n = self._n(basis)
        out = torch.empty((len(ks), n, n), dtype=torch.complex64, device=self.device)
        _lib.check(_lib.load().bem_assemble_batch_k(self.mesh.handle, ks.ctypes.data, len(ks), _lib.OPS[layer_type],
                                                    _lib.BASES[basis], _lib.QUADS[quad], int(clamp), _ptr(out), n,
                                                    n * n, _stream()), "bem_assemble_batch_k")
        return out

    def identity_matrix(self, basis=None):
        basis = basis or self.basis
        n = self._n(basis)
        out = torch.empty((n, n), dtype=torch.complex64, device=self.device)
        _lib.check(_lib.load().bem_identity(self.mesh.handle, _lib.BASES[basis], _ptr(out), n, _stream()),
                   "bem_identity")
        return out

    def _solve(self, LHS, RHS, solver, tol, restart, maxiter):
        if solver == "bicgstab":
            x, info = bicgstab(LHS, RHS)
        else:
            x, info = gmres(LHS, RHS, tol=tol, restart=restart, maxiter=maxiter)
        self.last_solve_info = info
        return x

    def CBIE(self, k, neumann, solver="gmres", tol=1e-6, restart=100, maxiter=1000):
        """Conventional BIE (solver.py:97-104): (M/2 - K) phi = -V psi.  For
        DP0 the system is built in one fused pass (V psi without storing V)."""
        if self.basis == "DP0":
            LHS, RHS = self.cbie_system(k, neumann)
            return self._solve(LHS, RHS, solver, tol, restart, maxiter)
        slp = self.assemble_boundary_matrix(k, "single")
        RHS = matvec(slp, neumann.to(torch.complex64).contiguous(), alpha=-1.0)
        del slp
        LHS = self.assemble_boundary_matrix(k, "double")
        scale_(LHS, -1.0)
        add_identity_(self, LHS, 0.5)
        return self._solve(LHS, RHS, solver, tol, restart, maxiter)

    def cbie_system(self, k, neumann, approx=False, out=None):
        """(M/2 - K, -V psi) for the DP0 basis in one fused pass."""
        if self.basis != "DP0":
            raise ValueError("the fused CBIE system is built for the DP0 basis")
        n = self.mesh.nt
        psi = neumann.to(torch.complex64).contiguous()
        lhs = torch.empty((n, n), dtype=torch.complex64, device=self.device) if out is None else out
        rhs = torch.empty(n, dtype=torch.complex64, device=self.device)
        lib = _lib.load()
        work = torch.empty(lib.bem_cbie_work_size(self.mesh.handle), dtype=torch.uint8, device=self.device)
        quad = "approx_1" if approx == "approx_1" else ("approx" if approx else "full")
        _lib.check(lib.bem_cbie_dp0(self.mesh.handle, float(k), _lib.QUADS[quad], int(self.clamp), _ptr(psi),
                                    _ptr(lhs), lhs.stride(0), _ptr(rhs), _ptr(work), _stream()), "bem_cbie_dp0")
        return lhs, rhs

    def HBIE(self, k, neumann, solver="gmres", tol=1e-6, restart=100, maxiter=1000):
        """Burton-Miller combined BIE (solver.py:106-116), beta = i/k:
        (M/2 - K + beta W) phi = (-V - beta (K' + M/2)) psi."""
        beta = 1j / k
        identity = self.identity_matrix()
        dlp = self.assemble_boundary_matrix(k, "double")
        hyp = self.assemble_boundary_matrix(k, "hypersingular")
        LHS = combine(combine(identity, 0.5, dlp, -1.0), 1.0, hyp, beta)
        del dlp, hyp
        slp = self.assemble_boundary_matrix(k, "single")
        adlp = self.assemble_boundary_matrix(k, "adjointdouble")
        R = combine(combine(slp, -1.0, adlp, -beta), 1.0, identity, -0.5 * beta)
        del slp, adlp
        RHS = matvec(R, neumann.to(torch.complex64).contiguous())
        return self._solve(LHS, RHS, solver, tol, restart, maxiter)

    def burton_miller_system(self, k, neumann, approx=True):
        """The system neumann2dirichlet solves (solver.py:120-125), with the
        approximate rules: (bm_lhs + M/2, (bm_rhs - (i/k) M/2) psi).  One
        library call (bem_burton_miller_system): the right-hand side's operator
        is contracted with psi inside the LHS kernels, never stored, and the
        mass terms are added in place (no dense identity)."""
        if self.basis != "P1":
            raise ValueError("the Burton-Miller system is built for the P1 basis")
        quad = "approx_1" if approx == "approx_1" else ("approx" if approx else "full")
        n = self._n("P1")
        g = neumann.to(torch.complex64).contiguous()
        check_tensor(g, torch.complex64)
        LHS = torch.empty((n, n), dtype=torch.complex64, device=self.device)
        RHS = torch.empty(n, dtype=torch.complex64, device=self.device)
        _lib.check(_lib.load().bem_burton_miller_system(self.mesh.handle, float(k), _lib.QUADS[quad],
                                                        int(self.clamp), _ptr(g), _ptr(LHS), n, _ptr(RHS),
                                                        _stream()), "bem_burton_miller_system")
        return LHS, RHS

    def neumann2dirichlet(self, k, neumann, solver="gmres", tol=1e-6, restart=100, maxiter=1000):
        """Burton-Miller with the approximate rules (solver.py:118-126):
        (bm_lhs + M/2) phi = (bm_rhs - beta M/2) psi."""
        LHS, RHS = self.burton_miller_system(k, neumann)
        return self._solve(LHS, RHS, solver, tol, restart, maxiter)

    def burton_miller_systems(self, ks, neumanns):
        """The systems neumann2dirichlet solves, at every wavenumber of ``ks``
        (``neumanns``: [K, n]): lhs [K, n, n] = bm_lhs + M/2, rhs [K, n] =
        (bm_rhs - (i/k) M/2) psi, all wavenumbers in one launch per kernel
        (bem_burton_miller_systems: the right-hand side's operator contracted
        with psi inside the LHS kernels, never stored)."""
        if self.basis != "P1":
            raise ValueError("the Burton-Miller system is built for the P1 basis")
        kh = np.ascontiguousarray(np.asarray([float(k) for k in ks], dtype=np.float32))
        neumanns = torch.as_tensor(neumanns, device=self.device).to(torch.complex64).reshape(len(kh), -1)
        neumanns = neumanns.contiguous()
        n = self._n("P1")
        lhs = torch.empty((len(kh), n, n), dtype=torch.complex64, device=self.device)
        rhs = torch.empty((len(kh), n), dtype=torch.complex64, device=self.device)
        _lib.check(_lib.load().bem_burton_miller_systems(self.mesh.handle, kh.ctypes.data, len(kh),
                                                         _lib.QUADS["approx"], int(self.clamp), _ptr(neumanns),
                                                         _ptr(lhs), n, n * n, _ptr(rhs), _stream()),
                   "bem_burton_miller_systems")
        return lhs, rhs

    def neumann2dirichlet_batch(self, ks, neumanns, tol=1e-6, restart=100, max_cycles=10):
        """neumann2dirichlet at every wavenumber of ``ks`` (the per-mode loop of
        scene.py:384-397 in one pass): the Burton-Miller systems of all
        wavenumbers (burton_miller_systems) solved by one lockstep batched
        GMRES.  Returns (dirichlet [K, n], infos)."""
        from .batch import batch_gmres

        lhs, rhs = self.burton_miller_systems(ks, neumanns)
        xs, infos = batch_gmres(list(lhs), list(rhs), tol=tol, restart=restart, max_cycles=max_cycles)
        self.last_solve_info = infos
        return torch.stack(xs), infos

    # -- potentials -----------------------------------------------------------
    def boundary2potential_batch(self, ks, neumanns, dirichlets, points):
        """boundary2potential at every wavenumber of ``ks``: p [K, P] = D phi -
        S psi, the potential operators of all wavenumbers in one launch each."""
        ks = [float(k) for k in ks]
        neumanns = torch.as_tensor(neumanns, device=self.device).to(torch.complex64).reshape(len(ks), -1)
        dirichlets = torch.as_tensor(dirichlets, device=self.device).to(torch.complex64).reshape(len(ks), -1)
        neumanns, dirichlets = neumanns.contiguous(), dirichlets.contiguous()
        S = self.potential_matrix_batch(ks, points, "single")
        D = self.potential_matrix_batch(ks, points, "double")
        p = matvec_strided(D, dirichlets)
        return matvec_strided(S, neumanns, alpha=-1.0, beta=1.0, out=p)

    def potential_matrix(self, k, points, layer_type, basis=None, clamp=None):
        basis = basis or self.basis
        clamp = self.clamp if clamp is None else clamp
        points = points.to(torch.float32).contiguous()
        check_tensor(points, torch.float32)
        n = self._n(basis)
        out = torch.empty((points.shape[0], n), dtype=torch.complex64, device=self.device)
        _lib.check(_lib.load().bem_potential(self.mesh.handle, _ptr(points), points.shape[0], float(k),
                                             _lib.OPS[layer_type], _lib.BASES[basis], int(clamp), _ptr(out), n,
                                             _stream()), "bem_potential")
        return out

    def potential_matrix_dk(self, k, points, layer_type, basis=None, clamp=None):
        """d/dk of potential_matrix (single or double layer)."""
        basis = basis or self.basis
        clamp = self.clamp if clamp is None else clamp
        points = points.to(torch.float32).contiguous()
        check_tensor(points, torch.float32)
        n = self._n(basis)
        out = torch.empty((points.shape[0], n), dtype=torch.complex64, device=self.device)
        _lib.check(_lib.load().bem_potential_dk(self.mesh.handle, _ptr(points), points.shape[0], float(k),
                                                _lib.OPS[layer_type], _lib.BASES[basis], int(clamp), _ptr(out), n,
                                                _stream()), "bem_potential_dk")
        return out

    def potential_matrix_batch(self, wavenumbers, points, layer_type, basis=None, clamp=None):
        """potential_matrix at every wavenumber: complex64 [len, np, n]."""
        basis = basis or self.basis
        clamp = self.clamp if clamp is None else clamp
        points = points.to(torch.float32).contiguous()
        check_tensor(points, torch.float32)
        ks = np.ascontiguousarray(np.asarray(wavenumbers, dtype=np.float32).reshape(-1))
        n, npts = self._n(basis), points.shape[0]
        out = torch.empty((len(ks), npts, n), dtype=torch.complex64, device=self.device)
        _lib.check(_lib.load().bem_potential_batch_k(self.mesh.handle, _ptr(points), npts, ks.ctypes.data, len(ks),
                                                     _lib.OPS[layer_type], _lib.BASES[basis], int(clamp), _ptr(out), n,
                                                     npts * n, _stream()), "bem_potential_batch_k")
        return out

    def single_potential(self, k, neumann, points):
        return matvec(self.potential_matrix(k, points, "single"), neumann.to(torch.complex64).contiguous())

    def double_potential(self, k, dirichlet, points):
        return matvec(self.potential_matrix(k, points, "double"), dirichlet.to(torch.complex64).contiguous())

    def boundary2potential(self, k, neumann, dirichlet, points):
        """p = -S psi + D phi (solver.py:140-147)."""
        S = self.potential_matrix(k, points, "single")
        D = self.potential_matrix(k, points, "double")
        p = matvec(D, dirichlet.to(torch.complex64).contiguous())
        return matvec(S, neumann.to(torch.complex64).contiguous(), alpha=-1.0, beta=1.0, out=p)


def map_triangle2vertex(vertices, triangles, features):
    """Vertex average of per-triangle complex values (solver.py:150-154), on
    the triangles alone (no mesh handle is built)."""
    check_tensor(vertices, torch.float32)
    check_tensor(triangles, torch.int32)
    check_tensor(features, torch.complex64)
    nv, nt = vertices.shape[0], triangles.shape[0]
    if features.shape[0] != nt:
        raise ValueError("features must hold one value per triangle")
    lib = _lib.load()
    out = torch.empty(nv, dtype=torch.complex64, device=vertices.device)
    work = torch.empty(max(int(lib.bem_triangle2vertex_work_size(nv, nt)), 1), dtype=torch.uint8,
                       device=vertices.device)
    with torch.cuda.device(vertices.device):
        _lib.check(lib.bem_triangle2vertex_raw(_ptr(triangles), nt, nv, _ptr(features), _ptr(out), _ptr(work),
                                               _stream()), "bem_triangle2vertex_raw")
    return out
