#!/usr/bin/env python
"""Benchmark of the BEM Helmholtz hot path on one MI355X.

Workload (single large mesh): icosphere level 6 (81920 triangles) of radius 1
with a seeded spherical-harmonic radial perturbation (degree <= 8, amplitude
5 %), piecewise-constant (DP0) complex64 single- and double-layer operators at
k in {10, 30, 60} rad/m, Neumann data from 4 seeded interior monopoles,
CBIE (M/2 - K) phi = -V psi solved by device GMRES to a relative residual of
1e-6, then the exterior pressure at 64 points (radius 1.5-3) checked against
the analytic field of the monopoles.

One step = the three solves (assembly of K and V, GMRES, exterior potential).
``value`` = operator entries assembled (2 N^2 per wavenumber) per second of
the whole timed region; ``assembly_entries_per_s`` counts assembly time only.

The ``roofline`` object is for the dominant kernel, the complex64 GEMV inside
GMRES (over 90 % of the timed region): algorithmic bytes per launch
8 N^2 + 16 N, divided by its mean duration from HIP events recorded around
every product on the stream GMRES launches it on.  ``roofline_assembly`` is
the fused regular-pair DP0 kernel that evaluates the double layer (stored as
-K) and the single layer (contracted with psi): algorithmic FLOP per launch =
(29 + 16) FLOP x 36 Gauss-point pairs x (N^2 - neighbour pairs) over its
HIP-event duration.

Beside the plain method, the same mesh runs through the reference's own
production path (P1 Burton-Miller, neumann2dirichlet) at every k, and the
level-6 systems' rows are checked against FP64 oracle rows in the CPU leg
(with the FP64 residual of the GPU solutions over those rows).

The CPU leg (at the end of the run, on the host cores) is the only part that
runs the oracle: ``cpu_baseline`` times its FP32 mirror of the reference
arithmetic (OpenMP, all cores of this process's affinity) on a row block of
the same operators -- the reference's own CPU path (bempp-cl/numba) is not
installed -- and the FP64 oracle answers every GPU result the forms recorded
for comparison (``Fp64Checks``) as the checker.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "BEM entries assembled/s + solves/s at <=1e-3 rel err vs FP64 (plain vs batched)"
FP32_PEAK = 157.3e12  # MI355X FP32 vector/matrix peak, FLOP/s
HBM_PEAK = 8.0e12  # MI355X HBM3E, B/s
# HBM bytes per matrix entry per launch (FETCH_SIZE + WRITE_SIZE, scaled by
# the counters' reading for a copy of known size), from the hardware-counter
# passes over ``bench.py --traffic-probe`` at level 6 kept under profiles/.
GEMV_TRAFFIC_PER_ENTRY = 8.0007  # 53.69 GB fetched + 0.7 MB written per level-6 launch
CBIE_TRAFFIC_PER_ENTRY = 10.231  # 14.0 GB fetched (geometry records) + 54.7 GB written
DLP_FLOP_PER_EVAL = 29  # algorithmic count per Gauss-point pair (SLP: 16)
SLP_FLOP_PER_EVAL = 16


def _per_entry(x, n):
    return None if x is None else x * n * n


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=2)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--level", type=int, default=6)
    p.add_argument("--ks", type=str, default="10,30,60")
    p.add_argument("--tol", type=float, default=1e-6)
    p.add_argument("--restart", type=int, default=1500)
    p.add_argument("--maxiter", type=int, default=3000)
    p.add_argument("--cpu-seconds", type=float, default=15.0)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--check-level", type=int, default=3, help="mesh level of the FP64 oracle check (0 = skip)")
    p.add_argument("--no-forms", action="store_true", help="skip the precision / batching / differentiation forms")
    p.add_argument("--no-bm", action="store_true",
                   help="skip the Burton-Miller (P1 neumann2dirichlet) solves on the workload mesh")
    p.add_argument("--batch-meshes", type=int, default=64)
    p.add_argument("--batch-ks", type=int, default=8)
    p.add_argument("--traffic-probe", action="store_true",
                   help="run one fused assembly launch, one GEMV and a 4 GiB copy, for hardware-counter passes")
    p.add_argument("--batch-probe", action="store_true",
                   help="one persistent assembly, one lockstep complex64 GMRES and one lockstep fp16 mixed-precision "
                        "solve of the first 8 batching meshes x 8 k, for kernel traces")
    p.add_argument("--p1-probe", action="store_true",
                   help="assemble the level-5 P1 single- and double-layer operators once each, for counter passes")
    p.add_argument("--c64-probe", action="store_true",
                   help="the 16-right-hand-side complex64 matrix-core product on the assembled level-6 system, "
                        "timed, then a 4 GiB copy (kernel traces and counter passes)")
    p.add_argument("--read-probe", action="store_true",
                   help="the GEMV beside torch.mv and a torch sum over the assembled level-6 matrix, timed")
    p.add_argument("--only-form", type=str, default="",
                   help="run one form alone and print its JSON (p1, batching, batching_small_meshes, "
                        "differentiation, differentiation_burton_miller, datagen, modal, burton_miller_level6, precision)")
    p.add_argument("--modal-sound-probe", action="store_true",
                   help="three batched modal-sound solves (32 modes of the modal form's block), for kernel traces")
    return p.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_workload(level, seed=0):
    from neuralat_amd import mesh

    v, t = mesh.icosphere(level, 1.0)
    v = mesh.sh_perturb(v, degree=8, amplitude=0.05, seed=seed)
    rng = np.random.default_rng(seed + 1)
    d = rng.standard_normal((4, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    src = d * rng.uniform(0.0, 0.5, size=(4, 1))
    pts = mesh.random_exterior_points(64, 1.5, 3.0, seed=seed + 2)
    return v, t, src, pts


class Solver:
    """The plain-method CBIE pipeline on the C ABI (device resident)."""

    def __init__(self, v, t, clamp=True):
        import torch

        from neuralat_amd.bem.solver import Mesh

        self.torch = torch
        self.dev = torch.device("cuda:0")
        self.vt = torch.tensor(v.astype(np.float32), device=self.dev)
        self.tt = torch.tensor(t.astype(np.int32), device=self.dev)
        self.mesh = Mesh(self.vt, self.tt)
        self.n = self.mesh.nt
        self.clamp = int(clamp)
        from neuralat_amd import _lib

        self.lhs = torch.empty((self.n, self.n), dtype=torch.complex64, device=self.dev)
        self.rhs = torch.empty(self.n, dtype=torch.complex64, device=self.dev)
        self.work = torch.empty(_lib.load().bem_cbie_work_size(self.mesh.handle), dtype=torch.uint8, device=self.dev)
        self.reg_ms = []

    def assemble(self, k, psi):
        """lhs = M/2 - K, rhs = -V psi in one fused pass (V never stored)."""
        from neuralat_amd import _lib
        from neuralat_amd.bem.solver import _ptr, _stream

        torch = self.torch
        lib = _lib.load()
        st = _stream()
        n = self.n
        e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
        e0.record()
        _lib.check(lib.bem_cbie_dp0_ex(self.mesh.handle, float(k), 0, self.clamp, 1, _ptr(psi), _ptr(self.lhs), n,
                                       _ptr(self.rhs), _ptr(self.work), st))
        e1.record()
        _lib.check(lib.bem_cbie_dp0_ex(self.mesh.handle, float(k), 0, self.clamp, 2, _ptr(psi), _ptr(self.lhs), n,
                                       _ptr(self.rhs), _ptr(self.work), st))
        e2.record()
        self._events = (e0, e1, e2)

    def timings(self):
        e0, e1, e2 = self._events
        return e0.elapsed_time(e1), e0.elapsed_time(e2)

    def solve(self, k, psi, pts_t, tol, restart, maxiter):
        from neuralat_amd.bem.solver import gmres

        self.assemble(k, psi)
        phi, info = gmres(self.lhs, self.rhs, tol=tol, restart=restart, maxiter=maxiter, time_products=True)
        S = self.potential(k, pts_t, "single")
        D = self.potential(k, pts_t, "double")
        from neuralat_amd.bem.solver import matvec

        p = matvec(D, phi)
        matvec(S, psi, alpha=-1.0, beta=1.0, out=p)
        return phi, p, info

    def potential(self, k, pts_t, op):
        from neuralat_amd import _lib
        from neuralat_amd.bem.solver import _ptr, _stream

        out = self.torch.empty((pts_t.shape[0], self.n), dtype=self.torch.complex64, device=self.dev)
        _lib.check(_lib.load().bem_potential(self.mesh.handle, _ptr(pts_t), pts_t.shape[0], float(k),
                                             _lib.OPS[op], 1, self.clamp, _ptr(out), self.n, _stream()))
        return out


class Fp64Checks:
    """GPU results of DP0 CBIE solves recorded for comparison with the FP64
    oracle solution of the same discrete system; the CPU leg computes the
    oracle answers at the end of the run and writes the relative errors into
    the result dictionaries (mode "set", or "max" over the items added)."""

    def __init__(self):
        self.items = []

    def add(self, into, v, t, k, psi, phis, pts=None, ps=None, mode="set"):
        """phis / ps: {result key: GPU surface solution / pressure at pts}."""
        self.items.append({"kind": "cbie", "into": into, "v": np.asarray(v, dtype=np.float32), "t": t,
                           "k": float(k), "psi": np.asarray(psi), "phis": phis, "pts": pts, "ps": ps or {},
                           "mode": mode})
        for key in list(phis) + list(ps or {}):
            if mode == "max":
                into.setdefault(key, 0.0)
            else:
                into[key] = None

    def add_rows(self, into, kind, v, t, k, rows, a_rows, b_rows, x, psi, clamp=True):
        """Rows ``rows`` of a GPU system (a_rows [r, n], b_rows [r]) and its GPU
        solution x: the CPU leg writes the relative error of the rows against
        FP64 oracle rows and the FP64 residual ||b - A x|| / ||b|| over them.
        kind "dp0_cbie": (M/2 - K, -V psi); kind "p1_bm": (bm_lhs + M/2,
        (bm_rhs - (i/k) M/2) psi) with the approximate rules."""
        self.items.append({"kind": kind, "into": into, "v": np.asarray(v, dtype=np.float32), "t": t, "k": float(k),
                           "rows": np.asarray(rows), "a": a_rows, "b": b_rows, "x": x, "psi": np.asarray(psi),
                           "clamp": bool(clamp)})
        for key in ("rows_rel_err_vs_fp64", "rhs_rel_err_vs_fp64", "fp64_residual_over_rows"):
            into[key] = None

    def add_modes(self, into, v, t, material, k, eigenvalues):
        """eigenvalues: {result key: GPU lowest non-rigid eigenvalues of the
        solid (v, t)}; the error written is the largest relative one."""
        self.items.append({"kind": "modes", "into": into, "v": np.asarray(v, dtype=np.float32), "t": t,
                           "material": material, "k": int(k), "eigs": eigenvalues})
        for key in eigenvalues:
            into[key] = None


CHECKS = Fp64Checks()


def fp64_check(level, ks):
    """Same pipeline on a small mesh; the FP64 oracle solution it is compared
    with comes from the CPU leg."""
    from neuralat_amd import mesh

    v, t, src, pts = make_workload(level)
    v32 = v.astype(np.float32)
    out = {}
    import torch

    s = Solver(v32, t)
    for k in ks:
        val, dn = mesh.sources_field(src, None, k)
        psi = mesh.project_dp0(v32, t, lambda x, n: dn(x, n))
        psi_t = torch.tensor(psi.astype(np.complex64), device=s.dev)
        pts_t = torch.tensor(pts.astype(np.float32), device=s.dev)
        phi, p, info = s.solve(k, psi_t, pts_t, 1e-6, 200, 2000)
        torch.cuda.synchronize()
        out[str(k)] = {}
        CHECKS.add(out[str(k)], v32, t, k, psi, {"surface_rel_err": phi.cpu().numpy()}, pts,
                   {"exterior_rel_err": p.cpu().numpy()})
    del s
    return out


def matvec_roofline(s, reps=5):
    """HBM roofline of the dense complex GEMV (the GMRES matvec) on the
    assembled operator: 8 N^2 + 16 N bytes per launch over its HIP-event time."""
    import torch

    from neuralat_amd.bem.solver import matvec

    n = s.n
    g = torch.Generator(device=s.dev).manual_seed(3)
    x = torch.randn(n, dtype=torch.complex64, device=s.dev, generator=g)
    y = torch.empty_like(x)
    matvec(s.lhs, x, out=y)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(reps + 1)]
    ev[0].record()
    for r in range(reps):
        matvec(s.lhs, x, out=y)
        ev[r + 1].record()
    torch.cuda.synchronize()
    ms = [ev[r].elapsed_time(ev[r + 1]) for r in range(reps)]
    nbytes = 8 * n * n + 16 * n
    sec = float(np.mean(ms)) / 1e3
    return {"bound": "hbm", "kernel": "gemv_n_vec_kernel<4> (complex64 GEMV)", "achieved": nbytes / sec / 1e9,
            "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": nbytes / sec / HBM_PEAK, "traffic": None,
            "bytes_per_launch": nbytes, "mean_launch_ms": sec * 1e3}


def _monopole_psis(v, t, k, srcs):
    from neuralat_amd import mesh

    out = []
    for src in srcs:
        _, dn = mesh.sources_field([src], None, k)
        out.append(mesh.project_dp0(v, t, lambda x, nn: dn(x, nn)))
    return np.array(out, dtype=np.complex64)


def precision_form(s, v, t, k, pts, nrhs=16, seed=5, check_level=4, restart=1500, tol=1e-5):
    """16 right-hand sides on the assembled level-6 system at k: plain FP32
    (one GMRES solve per right-hand side, timed on one of them, with the fused
    assembly and potentials of the plain path) against 16-bit storage +
    matrix-core multi-RHS GMRES + FP32 refinement (K and V assembled once,
    -V Psi on the FP32 matrix cores).  Errors: each solution against the
    analytic field at 64 exterior points, the FP32 solution of the same system,
    and the FP64 oracle on a level-``check_level`` mesh."""
    import torch

    from neuralat_amd import _lib, mesh
    from neuralat_amd.bem.precision import LowPrecisionMatrix, gmres_multi, matmul, solve_ir
    from neuralat_amd.bem.solver import _ptr, _stream

    torch.cuda.synchronize()
    rng = np.random.default_rng(seed)
    d = rng.standard_normal((nrhs, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    srcs = d * rng.uniform(0.0, 0.5, size=(nrhs, 1))
    psi = torch.tensor(_monopole_psis(v, t, k, srcs), device=s.dev)
    pts_t = torch.tensor(pts.astype(np.float32), device=s.dev)
    fields = np.array([mesh.sources_field([q], None, k)[0](pts) for q in srcs])
    n = s.n
    lib = _lib.load()
    res = {"k": k, "nrhs": nrhs, "tol": tol}

    # plain FP32 path, one right-hand side (the same for each of the 16)
    t0 = time.perf_counter()
    phi0, p0, info0 = s.solve(k, psi[0], pts_t, tol, restart, 3000)
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    res["plain_fp32"] = {"seconds_per_solve": t_plain, "solves_per_s": 1.0 / t_plain,
                         "gmres_iterations": info0["iterations"],
                         "rel_err_vs_analytic": float(np.linalg.norm(p0.cpu().numpy() - fields[0])
                                                      / np.linalg.norm(fields[0]))}
    log(f"bench: precision plain FP32 {t_plain:.2f} s/solve, {info0['iterations']} its")
    # K (fused pass, lhs) and V assembled once; B = -V Psi on the FP32 matrix
    # cores; then all 16 systems together: FP32 lockstep GMRES on the complex64
    # matrix, or 16-bit storage + flexible FP32 refinement
    for form in ("fp32_multi", "bf16", "fp16"):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s.assemble(k, psi[0])
        Vm = torch.empty((n, n), dtype=torch.complex64, device=s.dev)
        _lib.check(lib.bem_assemble(s.mesh.handle, float(k), _lib.OPS["single"], _lib.BASES["DP0"], 0, s.clamp,
                                    _ptr(Vm), n, _stream()), "bem_assemble")
        if form == "fp32_multi":
            # B = -V Psi on the FP32 matrix cores (16 right-hand sides, column
            # split): HIP events around the second of two products (the first
            # warms the clocks and the workspace pool) on the launch stream
            B = matmul(Vm, psi, alpha=-1.0)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            B = matmul(Vm, psi, alpha=-1.0)
            e1.record()
            torch.cuda.synchronize()
            pms = e0.elapsed_time(e1)
            nb = 8 * n * n + 16 * nrhs * n
            res["roofline_fp32_matrix_core_product"] = {
                "bound": "hbm", "kernel": "gemm_c64_mfma3_kernel (complex64, 16 right-hand sides, column split)",
                "achieved": nb / (pms / 1e3) / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": nb / (pms / 1e3) / HBM_PEAK, "ms": pms, "bytes_per_launch": nb,
                "byte_count": "8 B per matrix entry + 16 B per right-hand-side entry (X read, Y written)"}
            log(f"bench: FP32 matrix-core product {pms:.2f} ms ({nb / pms / 1e6:.0f} GB/s)")
        else:
            B = matmul(Vm, psi, alpha=-1.0)
        del Vm
        if form == "fp32_multi":
            X, infos = gmres_multi(s.lhs, B, tol=tol, restart=600, maxiter=3000)
            info = {"outer": infos[0]["iterations"], "inner": 0, "rel_residual": [i["rel_residual"] for i in infos],
                    "converged": all(i["converged"] for i in infos)}
        else:
            L = LowPrecisionMatrix(s.lhs, form)
            X, info = solve_ir(s.lhs, L, B, tol=tol, inner_tol=1e-2, restart=100, max_inner=300, max_outer=40)
            del L
        Sp = s.potential(k, pts_t, "single")
        Dp = s.potential(k, pts_t, "double")
        P = matmul(Dp, X)
        matmul(Sp, psi, alpha=-1.0, beta=1.0, out=P)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        Pc = P.cpu().numpy()
        errs = [float(np.linalg.norm(Pc[q] - fields[q]) / np.linalg.norm(fields[q])) for q in range(nrhs)]
        x0 = X[0]
        rel_fp32 = float((torch.linalg.vector_norm(x0 - phi0) / torch.linalg.vector_norm(phi0)).item())
        res[form] = {"seconds": el, "solves_per_s": nrhs / el, "speedup_vs_plain": (nrhs / el) * t_plain,
                     "outer_iterations": info["outer"], "inner_iterations": info["inner"],
                     "max_rel_residual": float(max(info["rel_residual"])), "converged": bool(info["converged"]),
                     "surface_rel_diff_vs_fp32_solution": rel_fp32,
                     "max_rel_err_vs_analytic": float(max(errs))}
        log(f"bench: precision {form}: {el:.2f} s for {nrhs} solves ({nrhs / el:.2f}/s, "
            f"{(nrhs / el) * t_plain:.1f}x plain), outer {info['outer']}, inner {info['inner']}, "
            f"residual {max(info['rel_residual']):.1e}")
    if check_level:
        res["fp64_check"] = precision_fp64_check(check_level, k, srcs[:2], tol)
    return res


def precision_fp64_check(level, k, srcs, tol):
    """The 16-bit + refinement pipeline on a small mesh of the same family;
    compared with the FP64 oracle solution of the same discrete system in the
    CPU leg."""
    import torch

    from neuralat_amd.bem.precision import LowPrecisionMatrix, solve_ir

    v, t, _, pts = make_workload(level)
    v32 = v.astype(np.float32)
    s = Solver(v32, t)
    psis = _monopole_psis(v32, t, k, srcs)
    out = {"level": level}
    for dt in ("bf16", "fp16"):
        rhs = []
        for q in range(len(srcs)):
            s.assemble(k, torch.tensor(psis[q], device=s.dev))
            rhs.append(s.rhs.clone())
        L = LowPrecisionMatrix(s.lhs, dt)
        X, info = solve_ir(s.lhs, L, torch.stack(rhs), tol=tol, inner_tol=1e-2, max_inner=300, max_outer=40)
        Xc = X.cpu().numpy()
        out[dt] = {"converged": bool(info["converged"])}
        for q in range(len(srcs)):
            CHECKS.add(out[dt], v32, t, k, psis[q], {"max_surface_rel_err": Xc[q]}, mode="max")
    del s
    return out


def grad_form(ndir=64, h=1e-3, seed=0):
    """Level-4 differentiation workload: loss = sum |p|^2 at 32 listeners from
    one interior monopole at k = 25; adjoint gradient w.r.t. all vertices and
    k, backward/forward time, and central differences along ``ndir`` seeded
    directions scaled to a root-mean-square vertex displacement h (relative
    to the unit radius), at h and h/2 with their Richardson combination
    (4 D(h/2) - D(h)) / 3, which removes the O(h^2) truncation."""
    import torch

    from neuralat_amd import mesh
    from neuralat_amd.bem.grad import exterior_pressure, monopole_neumann_dp0

    v, t = mesh.icosphere(4, 1.0)
    v = mesh.sh_perturb(v, 4, 0.05, seed=seed).astype(np.float32)
    rng = np.random.default_rng(seed + 10)
    pts = mesh.random_exterior_points(32, 1.5, 3.0, seed=seed + 3).astype(np.float32)
    src = rng.uniform(-0.2, 0.2, size=3)
    dev = torch.device("cuda:0")
    tt = torch.tensor(t, device=dev)
    pts_t = torch.tensor(pts, device=dev)
    k0 = 25.0

    def loss(vt, k, tol=1e-7):
        psi = monopole_neumann_dp0(vt, tt, src, k)
        p = exterior_pressure(vt, tt, k, psi, pts_t, tol=tol, adjoint_tol=max(tol, 1e-6))
        return (p.real ** 2 + p.imag ** 2).sum()

    tf, tb = [], []
    for _ in range(3):
        vt = torch.tensor(v, device=dev, requires_grad=True)
        kt = torch.tensor(k0, device=dev, requires_grad=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        L = loss(vt, kt, tol=1e-6)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        L.backward()
        torch.cuda.synchronize()
        tf.append(t1 - t0)
        tb.append(time.perf_counter() - t1)
    vt = torch.tensor(v, device=dev, requires_grad=True)
    kt = torch.tensor(k0, device=dev, requires_grad=True)
    L = loss(vt, kt)
    L.backward()
    gv = vt.grad.double().cpu().numpy()
    dirs = rng.standard_normal((ndir,) + v.shape)
    dirs *= np.sqrt(v.shape[0]) / np.linalg.norm(dirs.reshape(ndir, -1), axis=1)[:, None, None]
    quot = {}
    with torch.no_grad():
        for hh in (h, 0.5 * h):
            ad, fd = [], []
            for d in dirs:
                dp = torch.tensor((v + hh * d).astype(np.float32), device=dev)
                dm = torch.tensor((v - hh * d).astype(np.float32), device=dev)
                fd.append((float(loss(dp, k0)) - float(loss(dm, k0))) / (2 * hh))
                ad.append(float(np.sum(gv * (dp - dm).double().cpu().numpy())) / (2 * hh))
            quot[hh] = (np.array(ad), np.array(fd))
        hk = 1e-5 * k0
        fdk = (float(loss(vt.detach(), k0 + hk)) - float(loss(vt.detach(), k0 - hk))) / (2 * hk)
    (ad, fd), (ad2, fd2) = quot[h], quot[0.5 * h]
    adr, fdr = (4 * ad2 - ad) / 3, (4 * fd2 - fd) / 3
    return {"workload": "icosphere L4 (5120 triangles) + SH noise, k=25, 32 listeners, loss sum |p|^2",
            "forward_ms": 1e3 * min(tf), "backward_ms": 1e3 * min(tb), "backward_over_forward": min(tb) / min(tf),
            "fd_directions": ndir, "fd_step_rms_per_vertex": h,
            "fd_vertex_rel_diff": float(np.linalg.norm(ad - fd) / np.linalg.norm(fd)),
            "fd_vertex_rel_diff_richardson": float(np.linalg.norm(adr - fdr) / np.linalg.norm(fdr)),
            "dL_dk_adjoint": float(kt.grad), "dL_dk_fd": fdk}


def grad_bm_form(level=4, k=20.0, reps=3, seed=0):
    """The production path's gradient: the P1 Burton-Miller pressure
    (neumann2dirichlet + boundary2potential, approximate rules) at 32
    listeners on a level-``level`` sphere from one interior monopole; loss
    sum |p|^2, gradients with respect to every vertex, k, the Neumann data and
    the listener positions; forward and backward wall time (best of
    ``reps``).  Parity against the FP64 oracle is in tests/test_gpu_grad_bm.py."""
    import torch

    from neuralat_amd import mesh
    from neuralat_amd.bem.grad import burton_miller_pressure

    v, t = mesh.icosphere(level, 1.0)
    v = mesh.sh_perturb(v, 4, 0.05, seed=seed).astype(np.float32)
    _, dn = mesh.sources_field(np.array([[0.1, -0.05, 0.02]]), None, k)
    psi = mesh.project_p1(v, t, lambda x, n: dn(x, n)).astype(np.complex64)
    pts = mesh.random_exterior_points(32, 1.5, 3.0, seed=seed + 3).astype(np.float32)
    dev = torch.device("cuda:0")
    tt = torch.tensor(t, device=dev)
    tf, tb = [], []
    for _ in range(reps):
        vt = torch.tensor(v, device=dev, requires_grad=True)
        kt = torch.tensor(k, device=dev, requires_grad=True)
        gt = torch.tensor(psi, device=dev, requires_grad=True)
        pt = torch.tensor(pts, device=dev, requires_grad=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        p = burton_miller_pressure(vt, tt, kt, gt, pt)
        L = (p.abs() ** 2).sum()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        L.backward()
        torch.cuda.synchronize()
        tf.append(t1 - t0)
        tb.append(time.perf_counter() - t1)
    return {"workload": f"P1 Burton-Miller, icosphere L{level} ({v.shape[0]} vertices) + SH noise, k={k}, "
                        "32 listeners, loss sum |p|^2, gradients w.r.t. vertices, k, Neumann data, listeners",
            "forward_ms": 1e3 * min(tf), "backward_ms": 1e3 * min(tb), "backward_over_forward": min(tb) / min(tf),
            "dL_dk": float(kt.grad), "grad_vertices_norm": float(vt.grad.norm())}


def p1_form(level=5, k=10.0, bm_level=4, bm_k=20.0, reps=3):
    """The reference's own basis (P1, vertex unknowns): single- and
    double-layer assembly at ``level`` (HIP-event time per operator, entries/s
    and algorithmic FLOP/s: 36 Gauss-point pairs x (16 or 29 + 12 P1 weighting)
    FLOP per regular pair), and the Burton-Miller solve of the reference's data
    generation (neumann2dirichlet: bm_lhs / bm_rhs with the approximate rules,
    GMRES, boundary2potential) at ``bm_level`` against the analytic field."""
    import torch

    from neuralat_amd import mesh
    from neuralat_amd.bem.solver import BEM_Solver

    dev = torch.device("cuda:0")
    v, t = mesh.icosphere(level, 1.0)
    v = mesh.sh_perturb(v, 8, 0.05, seed=3).astype(np.float32)
    s = BEM_Solver(torch.tensor(v, device=dev), torch.tensor(t, device=dev), basis="P1")
    nv, npairs = v.shape[0], s.mesh.n_neighbor_pairs
    ntri = t.shape[0]
    out = {"assembly_level": level, "vertices": nv, "k": k}
    for op, fl in (("single", 16), ("double", 29)):
        s.assemble_boundary_matrix(k, op)
        torch.cuda.synchronize()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(reps + 1)]
        ev[0].record()
        for r in range(reps):
            A = s.assemble_boundary_matrix(k, op)
            ev[r + 1].record()
            del A
        torch.cuda.synchronize()
        ms = min(ev[r].elapsed_time(ev[r + 1]) for r in range(reps))
        flop = 36 * (fl + 12) * (ntri * ntri - npairs)
        out[op] = {"ms": ms, "entries_per_s": nv * nv / (ms / 1e3), "tflops_algorithmic": flop / (ms / 1e3) / 1e12}
        log(f"bench: P1 {op} layer level {level}: {ms:.1f} ms, {flop / (ms / 1e3) / 1e12:.1f} TFLOP/s algorithmic")
    # the Burton-Miller operators of the data generation (approximate rules:
    # 3-point triangle rule, 9 Gauss-point pairs per regular pair): the
    # regular-pair kernel alone (HIP events around bem_assemble_ex parts = 1)
    # and the whole operator
    from neuralat_amd import _lib
    from neuralat_amd.bem.solver import _ptr, _stream

    lib = _lib.load()
    A = torch.empty((nv, nv), dtype=torch.complex64, device=dev)
    for op in ("bm_lhs", "bm_rhs"):
        ms_reg, ms_all = [], []
        for r in range(reps + 1):
            e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
            e0.record()
            _lib.check(lib.bem_assemble_ex(s.mesh.handle, float(k), _lib.OPS[op], 0, 1, 1, 1, _ptr(A), nv,
                                           _stream()))
            e1.record()
            _lib.check(lib.bem_assemble_ex(s.mesh.handle, float(k), _lib.OPS[op], 0, 1, 1, 3, _ptr(A), nv,
                                           _stream()))
            e2.record()
            torch.cuda.synchronize()
            if r:
                ms_reg.append(e0.elapsed_time(e1))
                ms_all.append(e1.elapsed_time(e2))
        evals = 9 * (ntri * ntri - npairs)  # Gauss-point pairs of the regular pairs
        out[op] = {"rule": "approx (3-point triangle, 2-point line)", "regular_ms": min(ms_reg), "ms": min(ms_all),
                   "entries_per_s": nv * nv / (min(ms_all) / 1e3),
                   "regular_gauss_point_pairs_per_s": evals / (min(ms_reg) / 1e3)}
        log(f"bench: P1 {op} level {level}: regular pairs {min(ms_reg):.2f} ms, whole operator {min(ms_all):.2f} ms")
    del A
    del s
    torch.cuda.empty_cache()
    v, t = mesh.icosphere(bm_level, 1.0)
    v = mesh.sh_perturb(v, 8, 0.05, seed=4).astype(np.float32)
    val, dn = mesh.sources_field([[0.1, -0.05, 0.02]], None, bm_k)
    psi = torch.tensor(mesh.project_p1(v, t, lambda x, nn: dn(x, nn)).astype(np.complex64), device=dev)
    pts = mesh.random_exterior_points(64, 1.5, 3.0, seed=5)
    pts_t = torch.tensor(pts.astype(np.float32), device=dev)
    s = BEM_Solver(torch.tensor(v, device=dev), torch.tensor(t, device=dev), basis="P1")
    s.neumann2dirichlet(bm_k, psi)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    phi = s.neumann2dirichlet(bm_k, psi)
    p = s.boundary2potential(bm_k, psi, phi, pts_t)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    pe = val(pts)
    out["burton_miller"] = {"level": bm_level, "vertices": v.shape[0], "k": bm_k, "seconds": el,
                            "gmres_iterations": s.last_solve_info["iterations"],
                            "rel_err_vs_analytic": float(np.linalg.norm(p.cpu().numpy() - pe) / np.linalg.norm(pe))}
    log(f"bench: P1 Burton-Miller level {bm_level}: {el * 1e3:.1f} ms, "
        f"{s.last_solve_info['iterations']} GMRES its, error {out['burton_miller']['rel_err_vs_analytic']:.2e}")
    return out


def burton_miller_large_form(v, t, src, pts, ks, restart=1500, tol=1e-6, nrows=16):
    """The reference's production path on the workload mesh itself:
    neumann2dirichlet (P1 Burton-Miller, bm_lhs / bm_rhs with the approximate
    rules, src/bem/solver.py:118-126) + GMRES + boundary2potential at every k,
    with the r-clamp on (the reference's default) and off.  Per solve: wall
    time (assembly of both operators + mass terms + right-hand side, GMRES,
    exterior pressure), entries (2 V^2) per second, the error against the
    analytic field of the 4 monopoles; with the clamp on, rows of the system
    and the solution are kept for the FP64 row check of the CPU leg."""
    import torch

    from neuralat_amd import mesh
    from neuralat_amd.bem.solver import BEM_Solver, gmres

    dev = torch.device("cuda:0")
    v32 = v.astype(np.float32)
    s = BEM_Solver(torch.tensor(v32, device=dev), torch.tensor(t, device=dev), basis="P1")
    nv = v.shape[0]
    pts_t = torch.tensor(pts.astype(np.float32), device=dev)
    rng = np.random.default_rng(17)
    rows = np.sort(np.concatenate([np.arange(8), rng.choice(np.arange(8, nv), nrows - 8, replace=False)]))
    ridx = torch.tensor(rows, device=dev)
    psis, fields = {}, {}
    for k in ks:
        val, dn = mesh.sources_field(src, None, k)
        psis[k] = torch.tensor(mesh.project_p1(v32, t, lambda x, nn: dn(x, nn)).astype(np.complex64), device=dev)
        fields[k] = val(pts)
    out = {"vertices": nv, "basis": "P1", "entries_per_solve": 2 * nv * nv, "gmres_tol": tol,
           "gmres_restart": restart, "system": "(bm_lhs + M/2) phi = (bm_rhs - (i/k) M/2) psi, approximate rules"}
    lhs, rhs = s.burton_miller_system(ks[0], psis[ks[0]])  # warm-up
    del lhs, rhs
    for clamp in (True, False):
        s.clamp = clamp
        key = "clamp_on" if clamp else "clamp_off"
        out[key] = {}
        for k in ks:
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            t0 = time.perf_counter()
            e0.record()
            lhs, rhs = s.burton_miller_system(k, psis[k])
            e1.record()
            phi, info = gmres(lhs, rhs, tol=tol, restart=restart, maxiter=3000)
            p = s.boundary2potential(k, psis[k], phi, pts_t)
            torch.cuda.synchronize()
            el = time.perf_counter() - t0
            asm_ms = e0.elapsed_time(e1)
            err = float(np.linalg.norm(p.cpu().numpy() - fields[k]) / np.linalg.norm(fields[k]))
            res = {"seconds": el, "solves_per_s": 1.0 / el, "entries_per_s": 2 * nv * nv / el,
                   "assembly_ms": asm_ms, "assembly_entries_per_s": 2 * nv * nv / (asm_ms / 1e3),
                   "gmres_iterations": info["iterations"], "converged": bool(info["converged"]),
                   "rel_err_vs_analytic": err}
            if clamp:
                CHECKS.add_rows(res, "p1_bm", v32, t, k, rows, lhs.index_select(0, ridx).cpu().numpy(),
                                rhs.index_select(0, ridx).cpu().numpy(), phi.cpu().numpy(), psis[k].cpu().numpy())
            out[key][str(k)] = res
            log(f"bench: Burton-Miller level 6 ({key}) k={k}: {el:.2f} s (assembly {asm_ms:.0f} ms), "
                f"{info['iterations']} GMRES its, error vs analytic {err:.2e}")
            del lhs, rhs, phi
    del s
    return out


def datagen_form(n_samples=48, batch=None, level=2, seed=11):
    """The data-generation path (P1 Burton-Miller per sampled scene, |p| at
    64 targets): one neumann2dirichlet + boundary2potential per sample against
    the batched driver (the samples of a batch as one geometry set, their
    systems in one pass per kernel, one lockstep GMRES per batch; ``batch``
    None: the driver's default, here all 48 samples in one batch).
    Scene: two level-``level`` spheres, one rotating and resized, one moving."""
    import torch

    from neuralat_amd import mesh
    from neuralat_amd.bem.solver import BEM_Solver
    from neuralat_amd.datagen import SceneObject, SceneSampler

    v, t = mesh.icosphere(level, 0.1)
    objs = [SceneObject(v, t, resize=(1.0, 0.0, 0.0), rot_axis="z", rot_pos=(0.05, 0.0, 0.0), rot_max_deg=90.0,
                        vibration="x"),
            SceneObject(v + np.array([0.7, 0.0, 0.0]), t, move=(0.0, 0.2, 0.0), position=(0.0, 0.0, 0.1),
                        vibration="-z")]
    sc = SceneSampler(objs, trg_sample_num=64, freq_min=200.0, freq_max=4000.0, trg_pos_min=[-0.5] * 3,
                      trg_pos_max=[0.5] * 3)
    dev = torch.device("cuda:0")
    sc.generate(n_samples, seed=seed + 1, batch=batch)  # warm-up of the same shape (pool growth, first launches)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    x, y, infos = sc.generate(n_samples, seed=seed, batch=batch)
    torch.cuda.synchronize()
    t_batch = time.perf_counter() - t0
    rng = np.random.default_rng(seed)
    t0 = time.perf_counter()
    ref = []
    for q in range(n_samples):
        f = sc.sample_factors(rng)
        vv, tt, neu, k, pts = sc.build(f)
        s = BEM_Solver(torch.tensor(vv.astype(np.float32), device=dev), torch.tensor(tt, device=dev))
        psi = torch.tensor(neu, device=dev)
        phi = s.neumann2dirichlet(k, psi)
        ref.append(s.boundary2potential(k, psi, phi, torch.tensor(pts.astype(np.float32), device=dev)).abs())
    torch.cuda.synchronize()
    t_loop = time.perf_counter() - t0
    yr = torch.stack(ref).float().cpu()
    diff = float((torch.linalg.vector_norm(y[:, :, 0] - yr) / torch.linalg.vector_norm(yr)).item())
    log(f"bench: datagen {n_samples} samples: loop {t_loop:.2f} s, batched {t_batch:.2f} s")
    return {"workload": f"{n_samples} scenes of two level-{level} spheres (P1 Burton-Miller, 64 targets)",
            "samples_per_batch": int(min(n_samples, sc.batch_size() if batch is None else batch)),
            "vertices": int(2 * v.shape[0]), "loop_seconds": t_loop, "batched_seconds": t_batch,
            "speedup": t_loop / t_batch, "samples_per_s_batched": n_samples / t_batch,
            "converged": int(sum(i["converged"] for i in infos)), "rel_diff_batched_vs_loop": diff}


def make_batch(nmesh, nk, seed=7, levels=(2, 3, 4, 5)):
    """Seeded uneven meshes (icosphere levels taken in turn from ``levels``,
    anisotropic scale in [0.5, 2]^3, SH noise) and nk wavenumbers each,
    log-uniform in [5, 80]."""
    from neuralat_amd import mesh

    rng = np.random.default_rng(seed)
    out = []
    for q in range(nmesh):
        level = levels[q % len(levels)]
        v, t = mesh.icosphere(level, 1.0)
        v = mesh.anisotropic_scale(mesh.sh_perturb(v, 8, 0.05, seed=seed + q), rng.uniform(0.5, 2.0, 3))
        ks = np.exp(rng.uniform(np.log(5.0), np.log(80.0), size=nk))
        out.append((v.astype(np.float32), t, [float(x) for x in ks]))
    return out


def batch_form(nmesh=64, nk=8, restart=500, maxiter=2000, seed=7, levels=(2, 3, 4, 5), chunk_gb=115.0,
               mixed_dtype="fp16", fp64_items=8):
    """The batching workload (nmesh seeded meshes x nk wavenumbers), four ways:
    a plain per-item loop (fused assembly + GMRES per (mesh, k)); one
    persistent assembly launch + lockstep graph-captured complex64 GMRES; the
    same assembly + the lockstep mixed-precision solve (FP32 flexible GMRES,
    16-bit inner GMRES); the same with 8-bit (e4m3) inner copies.  The items'
    complex64 systems exceed one GPU's HBM for the full workload, so all four
    run on chunks of consecutive meshes
    whose systems fit ``chunk_gb`` (one persistent launch per chunk).  Neumann
    data: one interior monopole per mesh.  Also reports the matrix bytes the
    GMRES passes stream per second of solve time (sum over items of
    iterations x bytes per entry x n^2)."""
    import torch

    from neuralat_amd import mesh
    from neuralat_amd.bem.batch import (allocate_systems, batch_autotune, batch_cbie, batch_gmres,
                                        batch_solve_mixed)
    from neuralat_amd.bem.solver import BEM_Solver, gmres

    dev = torch.device("cuda:0")
    items = make_batch(nmesh, nk, seed, levels)
    chunks, cur, cur_b = [], [], 0.0
    for it in items:
        b = 8.0 * it[1].shape[0] ** 2 * len(it[2])
        if cur and cur_b + b > chunk_gb * 1e9:
            chunks.append(cur)
            cur, cur_b = [], 0.0
        cur.append(it)
        cur_b += b
    chunks.append(cur)
    max_cycles = max(1, maxiter // restart)
    acc = {"plain": 0.0, "batched": 0.0, "mixed": 0.0, "batched_asm": 0.0, "mixed_asm": 0.0, "mixed8": 0.0,
           "mixed8_asm": 0.0}
    its_plain, its_batch, its_mixed, its_mixed8, ntri_all = [], [], [], [], []
    conv = {"plain": 0, "batched": 0, "mixed": 0, "mixed8": 0}
    conv_all = 0  # items converged in every variant
    diffs_b, diffs_m, diffs_m8 = [], [], []
    fp64 = {}  # filled by the CPU leg
    nfp64 = 0
    best, tune_ms, done = None, None, 0
    for ci, chunk in enumerate(chunks):
        solvers, meshes, ks, psis, host = [], [], [], [], []
        for v, t, kl in chunk:
            sv = BEM_Solver(torch.tensor(v, device=dev), torch.tensor(t, device=dev), basis="DP0")
            solvers.append(sv)
            for k in kl:
                _, dn = mesh.sources_field([[0.05, -0.02, 0.01]], None, k)
                meshes.append(sv.mesh)
                ks.append(k)
                psis.append(torch.tensor(mesh.project_dp0(v, t, lambda x, nn: dn(x, nn)).astype(np.complex64),
                                         device=dev))
                host.append((v, t, k))
        outs = allocate_systems(meshes, dev)
        ntri = [m.nt for m in meshes]
        ntri_all += ntri
        # plain per-item loop
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        xs_plain, cp = [], []
        q = 0
        for sv, (v, t, kl) in zip(solvers, chunk):
            for k in kl:
                lhs, rhs = sv.cbie_system(k, psis[q], out=outs[q][0])
                x, info = gmres(lhs, rhs, tol=1e-6, restart=restart, maxiter=maxiter)
                xs_plain.append(x)
                conv["plain"] += int(info["converged"])
                cp.append(bool(info["converged"]))
                its_plain.append(info["iterations"])
                q += 1
        torch.cuda.synchronize()
        acc["plain"] += time.perf_counter() - t0
        if best is None:  # row tile for the persistent launch, tuned once (not timed)
            best, tune_ms = batch_autotune(meshes, ks, psis, outs)
        # persistent assembly + lockstep complex64 GMRES
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        batch_cbie(meshes, ks, psis, outs, row_tile=best)
        torch.cuda.synchronize()
        acc["batched_asm"] += time.perf_counter() - t0
        xs_b, inf_b = batch_gmres([o[0] for o in outs], [o[1] for o in outs], tol=1e-6, restart=restart,
                                  max_cycles=max_cycles, use_graph=True)
        torch.cuda.synchronize()
        acc["batched"] += time.perf_counter() - t0
        # persistent assembly + lockstep mixed-precision solve
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        batch_cbie(meshes, ks, psis, outs, row_tile=best)
        torch.cuda.synchronize()
        acc["mixed_asm"] += time.perf_counter() - t0
        xs_m, inf_m = batch_solve_mixed([o[0] for o in outs], [o[1] for o in outs], dtype=mixed_dtype, tol=1e-6,
                                        inner_tol=1e-3, outer_restart=10, inner_restart=restart,
                                        max_inner_cycles=1, max_outer_cycles=3)
        torch.cuda.synchronize()
        acc["mixed"] += time.perf_counter() - t0
        # persistent assembly + lockstep mixed-precision solve on 8-bit copies:
        # half the inner bytes, about 1 % from A, so a looser inner target and
        # one more outer cycle
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        batch_cbie(meshes, ks, psis, outs, row_tile=best)
        torch.cuda.synchronize()
        acc["mixed8_asm"] += time.perf_counter() - t0
        xs_m8, inf_m8 = batch_solve_mixed([o[0] for o in outs], [o[1] for o in outs], dtype="fp8", tol=1e-6,
                                          inner_tol=1e-2, outer_restart=10, inner_restart=restart,
                                          max_inner_cycles=1, max_outer_cycles=4)
        torch.cuda.synchronize()
        acc["mixed8"] += time.perf_counter() - t0
        its_batch += [i["iterations"] for i in inf_b]
        its_mixed += [i["inner_iterations"] for i in inf_m]
        its_mixed8 += [i["inner_iterations"] for i in inf_m8]
        cb = [bool(i["converged"]) for i in inf_b]
        cm = [bool(i["converged"]) for i in inf_m]
        cm8 = [bool(i["converged"]) for i in inf_m8]
        conv["batched"] += sum(cb)
        conv["mixed"] += sum(cm)
        conv["mixed8"] += sum(cm8)
        conv_all += sum(a and b and c and d for a, b, c, d in zip(cp, cb, cm, cm8))
        for xb, xm, xm8, xp, c1, c2, c3 in zip(xs_b, xs_m, xs_m8, xs_plain, cb, cm, cm8):
            nrm = torch.linalg.vector_norm(xp)
            if c1:
                diffs_b.append(float((torch.linalg.vector_norm(xb - xp) / nrm).item()))
            if c2:
                diffs_m.append(float((torch.linalg.vector_norm(xm - xp) / nrm).item()))
            if c3:
                diffs_m8.append(float((torch.linalg.vector_norm(xm8 - xp) / nrm).item()))
        # the first level-2 items (320 triangles), for the FP64 oracle in the CPU leg
        for q, (v, t, k) in enumerate(host):
            if nfp64 >= fp64_items or t.shape[0] != 320 or not (cb[q] and cm[q] and cm8[q]):
                continue
            CHECKS.add(fp64, v, t, k, psis[q].cpu().numpy(),
                       {"max_surface_rel_err_vs_fp64_level2_batched": xs_b[q].cpu().numpy(),
                        "max_surface_rel_err_vs_fp64_level2_mixed": xs_m[q].cpu().numpy(),
                        "max_surface_rel_err_vs_fp64_level2_mixed_fp8": xs_m8[q].cpu().numpy()}, mode="max")
            nfp64 += 1
        done += len(ks)
        log(f"bench: batching chunk {ci + 1}/{len(chunks)} ({done} solves): plain {acc['plain']:.2f} s, "
            f"batched {acc['batched']:.2f} s, mixed {acc['mixed']:.2f} s, mixed fp8 {acc['mixed8']:.2f} s")
        del outs, xs_plain, xs_b, xs_m, xs_m8, solvers, meshes, psis
        torch.cuda.empty_cache()
    nsol = len(ntri_all)
    entries = float(sum(2 * n * n for n in ntri_all))
    bytes_plain = float(sum(8.0 * n * n * it for n, it in zip(ntri_all, its_plain)))
    bytes_batch = float(sum(8.0 * n * n * it for n, it in zip(ntri_all, its_batch)))
    bytes_mixed = float(sum(4.0 * n * n * it for n, it in zip(ntri_all, its_mixed)))
    bytes_mixed8 = float(sum(2.0 * n * n * it for n, it in zip(ntri_all, its_mixed8)))
    lv = "-".join(f"L{x}" for x in levels)
    tp, tb, tm, tm8 = acc["plain"], acc["batched"], acc["mixed"], acc["mixed8"]
    res = fp64  # the CPU leg fills its FP64 entries in this dictionary
    res["fp64_items"] = nfp64
    res.update({"workload": f"{nmesh} meshes (icosphere {lv}, anisotropic scale [0.5,2], SH noise) x {nk} k "
                        f"log-uniform [5,80] = {nsol} DP0 CBIE solves to 1e-6, GMRES({restart}) at most "
                        f"{maxiter} iterations, in {len(chunks)} chunk(s) of at most {chunk_gb:.0f} GB of systems",
            "chunks": len(chunks), "triangles": sorted(set(ntri_all)),
            "plain_seconds": tp, "batched_seconds": tb, "mixed_seconds": tm,
            "solves_per_s_counts": "converged solves only (each variant's own count) per second of the whole "
                                   "variant; speedups are ratios of these; time_ratio_* compare the wall times "
                                   "of the same items",
            "plain_solves_per_s": conv["plain"] / tp, "batched_solves_per_s": conv["batched"] / tb,
            "mixed_solves_per_s": conv["mixed"] / tm,
            "speedup": (conv["batched"] / tb) / (conv["plain"] / tp),
            "speedup_mixed": (conv["mixed"] / tm) / (conv["plain"] / tp),
            "time_ratio_batched": tp / tb, "time_ratio_mixed": tp / tm, "time_ratio_mixed_fp8": tp / tm8,
            "converged_in_every_variant": conv_all,
            "mixed_dtype": mixed_dtype, "mixed_solver": "FP32 flexible GMRES(10), inner 16-bit GMRES(restart) "
                                                      "to max(1e-3, tol |b| / |r|), diagonal in FP32",
            "batched_assembly_seconds": acc["batched_asm"], "mixed_assembly_seconds": acc["mixed_asm"],
            "entries_per_s_batched": entries / tb, "entries_per_s_mixed": entries / tm,
            "gmres_iterations_total_plain": int(sum(its_plain)),
            "gmres_iterations_total_batched": int(sum(its_batch)),
            "inner_iterations_total_mixed": int(sum(its_mixed)),
            "matrix_GBps_plain": bytes_plain / tp / 1e9,
            "matrix_GBps_batched_gmres": bytes_batch / (tb - acc["batched_asm"]) / 1e9,
            "matrix_GBps_mixed_inner": bytes_mixed / (tm - acc["mixed_asm"]) / 1e9,
            "converged_plain": conv["plain"], "converged_batched": conv["batched"],
            "converged_mixed": conv["mixed"],
            "autotune_row_tile": best, "autotune_ms": tune_ms,
            "max_rel_diff_batched_vs_plain": max(diffs_b) if diffs_b else None,
            "max_rel_diff_mixed_vs_plain": max(diffs_m) if diffs_m else None,
            "mixed_fp8_seconds": tm8, "mixed_fp8_solves_per_s": conv["mixed8"] / tm8,
            "speedup_mixed_fp8": (conv["mixed8"] / tm8) / (conv["plain"] / tp),
            "mixed_fp8_solver": "FP32 flexible GMRES(10), inner GMRES(restart) on OCP e4m3 copies with a "
                                "power-of-two scale per row to max(1e-2, tol |b| / |r|), diagonal in FP32",
            "mixed_fp8_assembly_seconds": acc["mixed8_asm"],
            "inner_iterations_total_mixed_fp8": int(sum(its_mixed8)),
            "matrix_GBps_mixed_fp8_inner": bytes_mixed8 / (tm8 - acc["mixed8_asm"]) / 1e9,
            "converged_mixed_fp8": conv["mixed8"],
            "max_rel_diff_mixed_fp8_vs_plain": max(diffs_m8) if diffs_m8 else None})
    return res


def batch_probe(nmesh=8, nk=8, restart=500):
    """The first ``nmesh`` batching meshes x ``nk`` wavenumbers: one persistent
    assembly launch, one lockstep complex64 GMRES, one lockstep fp16 and one
    lockstep fp8 mixed-precision solve (the settings of the batching form)."""
    import torch

    from neuralat_amd import mesh
    from neuralat_amd.bem.batch import allocate_systems, batch_cbie, batch_gmres, batch_solve_mixed
    from neuralat_amd.bem.solver import BEM_Solver

    dev = torch.device("cuda:0")
    meshes, ks, psis = [], [], []
    for v, t, kl in make_batch(nmesh, nk):
        sv = BEM_Solver(torch.tensor(v, device=dev), torch.tensor(t, device=dev), basis="DP0")
        for k in kl:
            _, dn = mesh.sources_field([[0.05, -0.02, 0.01]], None, k)
            meshes.append(sv.mesh)
            ks.append(k)
            psis.append(torch.tensor(mesh.project_dp0(v, t, lambda x, nn: dn(x, nn)).astype(np.complex64),
                                     device=dev))
    outs = allocate_systems(meshes, dev)
    batch_cbie(meshes, ks, psis, outs)
    As, bs = [o[0] for o in outs], [o[1] for o in outs]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, ib = batch_gmres(As, bs, tol=1e-6, restart=restart, max_cycles=max(1, 2000 // restart))
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    _, im = batch_solve_mixed(As, bs, dtype="fp16", tol=1e-6, inner_tol=1e-3, outer_restart=10, inner_restart=restart,
                              max_inner_cycles=1, max_outer_cycles=3)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    _, i8 = batch_solve_mixed(As, bs, dtype="fp8", tol=1e-6, inner_tol=1e-2, outer_restart=10, inner_restart=restart,
                              max_inner_cycles=1, max_outer_cycles=4)
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    log(f"bench: batch probe {len(ks)} items: complex64 GMRES {t1 - t0:.3f} s "
        f"({sum(i['iterations'] for i in ib)} its, {sum(i['converged'] for i in ib)} converged), "
        f"fp16 mixed {t2 - t1:.3f} s ({sum(i['inner_iterations'] for i in im)} inner its, "
        f"{sum(i['converged'] for i in im)} converged), fp8 mixed {t3 - t2:.3f} s "
        f"({sum(i['inner_iterations'] for i in i8)} inner its, {sum(i['converged'] for i in i8)} converged)")


def modal_form(cells=(20, 10, 10), size=(0.2, 0.1, 0.1), k=32, reps=20):
    """FEM modal analysis of a plastic block (the step upstream of the BEM):
    HIP assembly of the mass and stiffness block matrices, the block product,
    the k lowest modes by the block LOBPCG on the library's FP64 products and
    by torch.lobpcg (the reference's solver, FP32), and the per-triangle
    Neumann data of every mode.  The FP64 oracle eigenvalues come from the CPU
    leg."""
    import torch

    from neuralat_amd import mesh
    from neuralat_amd.modal import FEMmodel, LOBPCG_solver, Material, MatSet, VibrationObj, lobpcg

    v, t = mesh.tet_box(*cells, size, jitter=0.2, seed=1)
    mat = MatSet.Plastic
    res = {"workload": f"tetrahedral block {cells[0]}x{cells[1]}x{cells[2]} cells ({len(t)} tetrahedra, "
                       f"{len(v)} vertices, {3 * len(v)} dofs), {size} m, plastic, k = {k} modes"}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    f = FEMmodel(v, t, Material(mat))
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    f._blocks.clear()
    f.blocks("stiffness")
    f.blocks("mass")
    ev1.record()
    torch.cuda.synchronize()
    asm_ms = ev0.elapsed_time(ev1)
    # block product: FP64 matrix and 48 columns (the LOBPCG block), bytes =
    # 9 * 8 per block + 4 (column index) + one read of X and one write of Y
    X = torch.randn(f.ndof, 48, dtype=torch.float64, device="cuda")
    f.matmul("stiffness", X)
    ev0.record()
    for _ in range(reps):
        f.matmul("stiffness", X)
    ev1.record()
    torch.cuda.synchronize()
    spmm_ms = ev0.elapsed_time(ev1) / reps
    spmm_bytes = f.nblocks * (72 + 4) + 2 * X.numel() * 8
    t0 = time.perf_counter()
    w, x, info = lobpcg(f, k)
    torch.cuda.synchronize()
    t_lob = time.perf_counter() - t0
    t0 = time.perf_counter()
    wt, _ = LOBPCG_solver(f.stiffness_matrix, f.mass_matrix, k)
    torch.cuda.synchronize()
    t_torch = time.perf_counter() - t0
    sv, st = mesh.tet_boundary(v, t)
    obj = VibrationObj(v, t, sv, st)
    t0 = time.perf_counter()
    obj.modal_analysis(k=k, material=Material(mat))
    neu = np.stack([obj.get_triangle_neumann(q) for q in range(k)])
    t_modal = time.perf_counter() - t0
    # modal sound: every mode's Burton-Miller solve and pressure at 256 points,
    # one mode at a time (the reference's loop) and all modes in lockstep
    from neuralat_amd.datagen import EditableModalSound, modal_preprocess

    obj.normalize(0.15)
    data = modal_preprocess(obj, k)
    ms = EditableModalSound(data, {"vibration_obj": {"size": 0.15},
                                   "sample": {"freq_rate": 1.0, "size_rate": 2.0, "bbox_rate": 2.0, "sample_num": 1,
                                              "point_num_per_sample": 256}})
    ms.sample(freqK_base=torch.tensor([1.0], device="cuda"), sizeK_base=torch.tensor([0.5], device="cuda"))
    ms.solve()  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    p_loop = ms.solve_loop()
    torch.cuda.synchronize()
    t_ms_loop = time.perf_counter() - t0
    t0 = time.perf_counter()
    p_batch = ms.solve()
    torch.cuda.synchronize()
    t_ms_batch = time.perf_counter() - t0
    res["modal_sound"] = {"modes": k, "surface_vertices": int(len(sv)),
                          "k_range": [float(ms.ks.abs().min()), float(ms.ks.abs().max())],
                          "loop_seconds": t_ms_loop, "batched_seconds": t_ms_batch, "speedup": t_ms_loop / t_ms_batch,
                          "converged": int(sum(i["converged"] for i in ms.solve_infos)),
                          "rel_diff_batched_vs_loop": float((torch.linalg.vector_norm(p_batch - p_loop)
                                                             / torch.linalg.vector_norm(p_loop)).item())}
    log(f"bench: modal sound {k} modes: loop {t_ms_loop:.3f} s, batched {t_ms_batch:.3f} s")
    res.update({"setup_seconds": t_setup, "assembly_ms": asm_ms, "tets_per_s": len(t) / (asm_ms / 1e3),
                "block_product_ms": spmm_ms, "block_product_GBps": spmm_bytes / (spmm_ms / 1e3) / 1e9,
                "lobpcg_seconds": t_lob, "lobpcg_iterations": info["iterations"],
                "lobpcg_max_rel_residual": info["max_rel_residual"], "torch_lobpcg_seconds": t_torch,
                "max_rel_diff_torch_vs_lobpcg": float(np.max(np.abs(wt - w) / w)),
                "modal_analysis_seconds": t_modal, "surface_triangles": int(len(st)),
                "frequencies_hz": [float(x) for x in (np.sqrt(w[:4]) / (2 * np.pi))],
                "neumann_finite": bool(np.isfinite(neu).all())})
    CHECKS.add_modes(res, v, t, mat, k, {"max_eig_rel_err_vs_fp64_lobpcg": w,
                                         "max_eig_rel_err_vs_fp64_torch_lobpcg": wt})
    log(f"bench: modal {len(t)} tets: assembly {asm_ms:.2f} ms, block product {spmm_ms * 1e3:.0f} us, "
        f"lobpcg {t_lob:.2f} s ({info['iterations']} its), torch.lobpcg {t_torch:.2f} s")
    return res


def fp64_answers(checks):
    """The FP64 oracle solution of each recorded system, (M/2 - K) phi = -V psi
    and p = D_pot phi - S_pot psi, against the recorded GPU results."""
    import oracle
    from oracle import fem as ofem

    for it in checks.items:
        if it["kind"] == "modes":
            rho, E, nu = it["material"][:3]
            K = ofem.assemble(it["v"], it["t"], "stiffness", youngs=E, poisson=nu)
            M = ofem.assemble(it["v"], it["t"], "mass", density=rho)
            w64, _ = ofem.modes(K, M, it["k"])
            for key, w in it["eigs"].items():
                it["into"][key] = float(np.max(np.abs(np.asarray(w) - w64) / w64))
            continue
        if it["kind"] in ("dp0_cbie", "p1_bm"):
            _fp64_rows_answer(it)
            continue
        v, t, k, psi = it["v"], it["t"], it["k"], it["psi"].astype(np.complex128)
        S = oracle.assemble(v, t, k, "single", "DP0")
        D = oracle.assemble(v, t, k, "double", "DP0")
        M = oracle.identity(v, t, "DP0")
        phi64 = np.linalg.solve(0.5 * M - D, -S @ psi)
        errs = {key: float(np.linalg.norm(x - phi64) / np.linalg.norm(phi64)) for key, x in it["phis"].items()}
        if it["ps"]:
            pts = it["pts"]
            p64 = -oracle.potential(v, t, pts, k, "single", "DP0") @ psi + oracle.potential(v, t, pts, k, "double",
                                                                                             "DP0") @ phi64
            errs.update({key: float(np.linalg.norm(p - p64) / np.linalg.norm(p64)) for key, p in it["ps"].items()})
        for key, e in errs.items():
            it["into"][key] = max(it["into"][key] or 0.0, e) if it["mode"] == "max" else e


def _p1_mass_rows(v, t, rows):
    import scipy.sparse as sp

    v = np.asarray(v, np.float64)
    area = 0.5 * np.linalg.norm(np.cross(v[t[:, 1]] - v[t[:, 0]], v[t[:, 2]] - v[t[:, 0]]), axis=1)
    r = np.concatenate([t[:, a] for a in range(3) for b in range(3)])
    c = np.concatenate([t[:, b] for a in range(3) for b in range(3)])
    d = np.concatenate([area / 12.0 * (1.0 + (a == b)) for a in range(3) for b in range(3)])
    return sp.csr_matrix((d, (r, c)), shape=(len(v), len(v)))[np.asarray(rows)].toarray()


def _fp64_rows_answer(it):
    """FP64 oracle rows of a recorded level-6 system and the residual of the
    GPU solution over them (oracle.entries / oracle.p1_rows: no full matrix)."""
    import oracle

    v, t, k, rows, psi = it["v"], it["t"], it["k"], it["rows"], it["psi"].astype(np.complex128)
    clamp = it["clamp"]
    if it["kind"] == "dp0_cbie":
        nt = t.shape[0]
        r = np.repeat(rows.astype(np.int32), nt)
        c = np.tile(np.arange(nt, dtype=np.int32), len(rows))
        K = oracle.entries(v, t, k, r, c, "double", clamp=clamp).reshape(len(rows), nt)
        V = oracle.entries(v, t, k, r, c, "single", clamp=clamp).reshape(len(rows), nt)
        v64 = v.astype(np.float64)
        area = 0.5 * np.linalg.norm(np.cross(v64[t[:, 1]] - v64[t[:, 0]], v64[t[:, 2]] - v64[t[:, 0]]), axis=1)
        A = -K
        A[np.arange(len(rows)), rows] += 0.5 * area[rows]
        b = -(V @ psi)
    else:
        M = _p1_mass_rows(v, t, rows)
        A = oracle.p1_rows(v, t, k, rows, "bm_lhs", quad="approx", clamp=clamp) + 0.5 * M
        b = (oracle.p1_rows(v, t, k, rows, "bm_rhs", quad="approx", clamp=clamp) - 0.5j / k * M) @ psi
    x = np.asarray(it["x"], np.complex128)
    into = it["into"]
    into["rows_rel_err_vs_fp64"] = float(np.linalg.norm(it["a"] - A) / np.linalg.norm(A))
    into["rhs_rel_err_vs_fp64"] = float(np.linalg.norm(it["b"] - b) / np.linalg.norm(b))
    into["fp64_residual_over_rows"] = float(np.linalg.norm(A @ x - b) / np.linalg.norm(b))
    into["rows_checked"] = int(len(rows))


def cpu_leg(args, v, t, k):
    """The CPU leg, on the host cores after every GPU measurement: the FP64
    oracle answers for the GPU results the run recorded, then (unless
    --no-cpu) the CPU baseline.  The only part of the benchmark that runs the
    oracle."""
    log(f"bench: CPU leg: FP64 oracle answers for {len(CHECKS.items)} recorded GPU results")
    fp64_answers(CHECKS)
    if args.no_cpu:
        return None
    log("bench: CPU leg: baseline")
    return cpu_baseline(v, t, k, args.cpu_seconds)


def cpu_baseline(v, t, k, seconds):
    """FP32 mirror of the reference arithmetic on rows [0, R) of the DP0 SLP
    and DLP operators, all cores of this process's affinity."""
    import oracle

    cores = len(os.sched_getaffinity(0))
    v32 = v.astype(np.float32)
    n = t.shape[0]
    rows = 8
    el = 0.0
    while True:
        t0 = time.perf_counter()
        oracle.assemble(v32, t, k, "single", "DP0", precision="f32", rows=(0, rows), nthreads=cores)
        oracle.assemble(v32, t, k, "double", "DP0", precision="f32", rows=(0, rows), nthreads=cores)
        el = time.perf_counter() - t0
        if el > seconds / 4 or rows >= n:
            break
        rows = min(n, int(rows * max(2.0, min(8.0, (seconds / 3) / max(el, 1e-3)))))
    rate = 2 * rows * n / el
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": rate, "unit": "entries/s", "cores": cores, "kind": "port",
            "sample": f"rows 0..{rows} of the level-{int(round(np.log(n / 20) / np.log(4)))} DP0 single+double layer "
                      f"operators ({2 * rows * n} entries, {el:.1f} s), FP32 mirror of the reference arithmetic, "
                      f"OpenMP; CPU: {model}"}


def traffic_probe(level, k):
    """One launch each of the fused assembly kernel and the GEMV on the
    level-``level`` workload, then a device copy of 4 GiB (read 4 GiB, write
    4 GiB) that calibrates the FETCH_SIZE / WRITE_SIZE counters."""
    import torch

    from neuralat_amd import mesh
    from neuralat_amd.bem.solver import matvec

    v, t, src, _ = make_workload(level)
    s = Solver(v.astype(np.float32), t)
    _, dn = mesh.sources_field(src, None, k)
    psi = torch.tensor(mesh.project_dp0(v, t, lambda x, nn: dn(x, nn)).astype(np.complex64), device=s.dev)
    s.assemble(k, psi)
    torch.cuda.synchronize()
    y = matvec(s.lhs, psi)
    torch.cuda.synchronize()
    a = torch.ones(1 << 30, dtype=torch.float32, device=s.dev)
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"triangles": s.n, "cbie_entries": s.n * s.n, "gemv_bytes": 8 * s.n * s.n + 16 * s.n,
                      "copy_read_bytes": 4 << 30, "copy_write_bytes": 4 << 30, "y0": str(complex(y[0].item()))}))


def read_probe(level, k, reps=6):
    """The GEMV's bandwidth beside two library reads of the same assembled
    level-``level`` CBIE matrix (complex64, 8 N^2 bytes), each timed with
    events, best of ``reps`` after one warm-up: this build's GEMV
    (``matvec``), torch's complex GEMV (``torch.mv``, the hipBLAS path that the
    reference's ``@`` takes) and torch's full sum over the matrix as real
    numbers (a plain read of every byte)."""
    import torch

    from neuralat_amd import mesh
    from neuralat_amd.bem.solver import matvec

    v, t, src, _ = make_workload(level)
    s = Solver(v.astype(np.float32), t)
    _, dn = mesh.sources_field(src, None, k)
    psi = torch.tensor(mesh.project_dp0(v, t, lambda x, nn: dn(x, nn)).astype(np.complex64), device=s.dev)
    s.assemble(k, psi)
    torch.cuda.synchronize()
    n = s.n
    nbytes = 8 * n * n
    A = s.lhs  # [n, n], contiguous

    def best(f):
        f()
        torch.cuda.synchronize()
        ms = []
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            f()
            e1.record()
            torch.cuda.synchronize()
            ms.append(e0.elapsed_time(e1))
        return min(ms)

    out = {"rows": n, "matrix_bytes": nbytes}
    y_ours = matvec(A, psi)
    y_torch = torch.mv(A, psi)
    out["torch_mv_rel_diff"] = float((torch.linalg.vector_norm(y_torch - y_ours) /
                                      torch.linalg.vector_norm(y_torch)).item())
    del y_ours, y_torch
    for name, f in (("gemv", lambda: matvec(A, psi)), ("torch_mv", lambda: torch.mv(A, psi)),
                    ("torch_sum", lambda: torch.view_as_real(A).sum())):
        ms = best(f)
        out[name] = {"ms": ms, "achieved_gb_s": nbytes / ms / 1e6, "frac_of_hbm_peak": nbytes / (ms / 1e3) / HBM_PEAK}
    print(json.dumps(out))


def c64_probe(level, k, nrhs=16, reps=6):
    """Y = A X on the FP32 matrix cores for the assembled level-``level``
    CBIE matrix A (complex64) and ``nrhs`` seeded right-hand sides, ``reps``
    launches timed with events (the product the refinement's residual and
    the potentials use), the first 512 rows checked against torch's product;
    then a device copy of 4 GiB for the FETCH_SIZE / WRITE_SIZE calibration."""
    import torch

    from neuralat_amd import mesh
    from neuralat_amd.bem.precision import matmul

    v, t, src, _ = make_workload(level)
    s = Solver(v.astype(np.float32), t)
    _, dn = mesh.sources_field(src, None, k)
    psi = torch.tensor(mesh.project_dp0(v, t, lambda x, nn: dn(x, nn)).astype(np.complex64), device=s.dev)
    s.assemble(k, psi)
    n = s.n
    g = torch.Generator(device=s.dev).manual_seed(1)
    X = torch.randn((nrhs, n), dtype=torch.complex64, device=s.dev, generator=g)
    ref = torch.matmul(s.lhs[:512], X.T.contiguous()).T
    algo = 8 * n * n + 16 * nrhs * n
    out = {"rows": n, "nrhs": nrhs, "algorithmic_bytes": algo}
    for split in (True, False):
        Y = torch.empty((nrhs, n), dtype=torch.complex64, device=s.dev)
        ms = []
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            matmul(s.lhs, X, out=Y, split=split)
            e1.record()
            torch.cuda.synchronize()
            ms.append(e0.elapsed_time(e1))
        err = float((torch.linalg.matrix_norm(Y[:, :512] - ref) / torch.linalg.matrix_norm(ref)).item())
        best = min(ms[1:])
        out["column_split" if split else "one_pass"] = {
            "ms": ms, "best_ms": best, "achieved_gb_s": algo / best / 1e6,
            "frac_of_hbm_peak": algo / (best / 1e3) / HBM_PEAK, "rel_err_vs_torch_first_512_rows": err}
        del Y
    a = torch.ones(1 << 30, dtype=torch.float32, device=s.dev)
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    out.update({"copy_read_bytes": 4 << 30, "copy_write_bytes": 4 << 30})
    print(json.dumps(out))


def modal_sound_probe(cells=(20, 10, 10), size=(0.2, 0.1, 0.1), k=32, reps=3):
    """The modal form's modal-sound step alone: the block's k lowest modes,
    then ``reps`` batched solves (every mode's Burton-Miller system and the
    pressure at 256 points), for kernel traces."""
    import torch

    from neuralat_amd import mesh
    from neuralat_amd.datagen import EditableModalSound, modal_preprocess
    from neuralat_amd.modal import Material, MatSet, VibrationObj

    v, t = mesh.tet_box(*cells, size, jitter=0.2, seed=1)
    sv, st = mesh.tet_boundary(v, t)
    obj = VibrationObj(v, t, sv, st)
    obj.modal_analysis(k=k, material=Material(MatSet.Plastic))
    obj.normalize(0.15)
    ms = EditableModalSound(modal_preprocess(obj, k), {"vibration_obj": {"size": 0.15},
                                                       "sample": {"freq_rate": 1.0, "size_rate": 2.0, "bbox_rate": 2.0,
                                                                  "sample_num": 1, "point_num_per_sample": 256}})
    ms.sample(freqK_base=torch.tensor([1.0], device="cuda"), sizeK_base=torch.tensor([0.5], device="cuda"))
    for q in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ms.solve()
        torch.cuda.synchronize()
        print(json.dumps({"solve": q, "modes": k, "surface_vertices": int(len(sv)),
                          "ms": 1e3 * (time.perf_counter() - t0)}))


def p1_probe(level=5, k=10.0):
    """One assembly of each P1 layer operator and of the two Burton-Miller
    operators with the approximate rules (the data-generation path) on the
    level-``level`` mesh of the P1 measurement, then neumann2dirichlet's
    system in one pass (for kernel traces and hardware-counter passes)."""
    import torch

    from neuralat_amd import mesh
    from neuralat_amd.bem.solver import BEM_Solver

    v, t = mesh.icosphere(level, 1.0)
    v = mesh.sh_perturb(v, 8, 0.05, seed=3).astype(np.float32)
    s = BEM_Solver(torch.tensor(v, device="cuda"), torch.tensor(t, device="cuda"), basis="P1")
    for op, approx in (("single", False), ("double", False), ("bm_lhs", True), ("bm_rhs", True)):
        A = s.assemble_boundary_matrix(k, op, approx=approx)
        torch.cuda.synchronize()
        print(json.dumps({"op": op, "approx": approx, "vertices": int(v.shape[0]),
                          "entry_0_5": str(complex(A[0, 5].item()))}))
        del A
    # neumann2dirichlet's system in one pass (bm_rhs contracted with g in the bm_lhs kernels)
    g = torch.ones(v.shape[0], dtype=torch.complex64, device="cuda")
    L, b = s.burton_miller_system(k, g)
    torch.cuda.synchronize()
    print(json.dumps({"op": "burton_miller_system", "vertices": int(v.shape[0]), "rhs_0": str(complex(b[0].item()))}))
    del L, b


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1",
                          "gpus_requested": args.gpus}))
        sys.exit(2)
    import torch

    from neuralat_amd import _lib, mesh

    _lib.load()
    ks = [float(x) for x in args.ks.split(",")]
    if args.traffic_probe:
        traffic_probe(args.level, ks[0])
        return
    if args.p1_probe:
        p1_probe()
        return
    if args.c64_probe:
        c64_probe(args.level, ks[-1])
        return
    if args.read_probe:
        read_probe(args.level, ks[-1])
        return
    if args.batch_probe:
        batch_probe()
        return
    if args.modal_sound_probe:
        modal_sound_probe()
        return
    if args.only_form:
        forms = {"p1": lambda: p1_form(), "batching": lambda: batch_form(args.batch_meshes, args.batch_ks),
                 "batching_small_meshes": lambda: batch_form(16, args.batch_ks, levels=(2, 3), seed=9),
                 "differentiation": lambda: grad_form(), "datagen": lambda: datagen_form(),
                 "differentiation_burton_miller": lambda: grad_bm_form(),
                 "modal": lambda: modal_form()}
        if args.only_form == "burton_miller_level6":
            v, t, src, pts = make_workload(args.level)
            res = burton_miller_large_form(v, t, src, pts, ks, restart=args.restart)
            fp64_answers(CHECKS)
        elif args.only_form == "precision":
            v, t, src, pts = make_workload(args.level)
            res = precision_form(Solver(v.astype(np.float32), t), v, t, ks[-1], pts, restart=args.restart)
            fp64_answers(CHECKS)
        else:
            res = forms[args.only_form]()
            fp64_answers(CHECKS)
        print(json.dumps({args.only_form: res}))
        return
    log(f"bench: building level-{args.level} workload")
    v, t, src, pts = make_workload(args.level)
    n = t.shape[0]
    fp64 = {}
    if args.check_level > 0:
        log(f"bench: level-{args.check_level} solves for the FP64 oracle check (answered in the CPU leg)")
        fp64 = fp64_check(args.check_level, ks)
    log(f"bench: allocating {n}x{n} complex64 ({8 * n * n / 1e9:.1f} GB)")
    s = Solver(v.astype(np.float32), t)
    dev = s.dev
    psis, fields = {}, {}
    for k in ks:
        val, dn = mesh.sources_field(src, None, k)
        psis[k] = torch.tensor(mesh.project_dp0(v, t, lambda x, nn: dn(x, nn)).astype(np.complex64), device=dev)
        fields[k] = val(pts)
    pts_t = torch.tensor(pts.astype(np.float32), device=dev)
    npairs = s.mesh.n_neighbor_pairs

    # rows of the level-6 system kept from the (untimed) warm-up for the FP64
    # row check of the CPU leg: the first 32 and 32 seeded others
    rows = np.sort(np.concatenate([np.arange(32), np.random.default_rng(5).choice(np.arange(32, n), 32,
                                                                                   replace=False)]))
    ridx = torch.tensor(rows, device=dev)
    fp64_rows = {}

    def step(record, capture=False):
        res = []
        for k in ks:
            phi, p, info = s.solve(k, psis[k], pts_t, args.tol, args.restart, args.maxiter)
            reg_ms, asm_ms = s.timings()
            res.append((k, p, info, reg_ms, asm_ms))
            if capture:
                fp64_rows[str(k)] = {}
                CHECKS.add_rows(fp64_rows[str(k)], "dp0_cbie", v, t, k, rows,
                                s.lhs.index_select(0, ridx).cpu().numpy(), s.rhs.index_select(0, ridx).cpu().numpy(),
                                phi.cpu().numpy(), psis[k].cpu().numpy(), clamp=bool(s.clamp))
            if record:
                log(f"bench: k={k} fused regular kernel {reg_ms:.1f} ms, assembly {asm_ms:.1f} ms, "
                    f"GMRES its {info['iterations']} res {info['rel_residual']:.2e}")
        return res

    for w in range(args.warmup):
        log(f"bench: warmup step {w}")
        step(True, capture=(w == 0))
    torch.cuda.synchronize()
    log(f"bench: timing {args.steps} steps")
    t0 = time.perf_counter()
    all_res = []
    for sidx in range(args.steps):
        all_res.append(step(True))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0

    reg = [r[3] for res in all_res for r in res]
    asm = [r[4] for res in all_res for r in res]
    gemv_ms = sum(r[2]["gemv_ms"] for res in all_res for r in res)
    gemv_n = sum(r[2]["gemv_count"] for res in all_res for r in res)
    gemv_bytes = 8 * n * n + 16 * n
    gemv_avg_s = gemv_ms / 1e3 / max(gemv_n, 1)
    errs, its = {}, {}
    for (k, p, info, _, _) in all_res[-1]:
        pe = fields[k]
        errs[str(k)] = float(np.linalg.norm(p.cpu().numpy() - pe) / np.linalg.norm(pe))
        its[str(k)] = info["iterations"]
    nsolves = args.steps * len(ks)
    entries = 2 * n * n * nsolves
    reg_mean_s = float(np.mean(reg)) / 1e3
    flop = (DLP_FLOP_PER_EVAL + SLP_FLOP_PER_EVAL) * 36 * (n * n - npairs)
    achieved = flop / reg_mean_s
    mv = matvec_roofline(s)
    log(f"bench: matvec {mv['achieved']:.0f} GB/s ({mv['frac']:.3f} of peak)")
    forms = {}
    if not args.no_forms:
        forms["precision"] = precision_form(s, v, t, ks[-1], pts, restart=args.restart)
    del s
    torch.cuda.empty_cache()
    if not args.no_bm:
        forms["burton_miller_level6"] = burton_miller_large_form(v, t, src, pts, ks, restart=args.restart)
        torch.cuda.empty_cache()
    if not args.no_forms:
        forms["batching"] = batch_form(args.batch_meshes, args.batch_ks)
        torch.cuda.empty_cache()
        forms["batching_small_meshes"] = batch_form(16, args.batch_ks, levels=(2, 3), seed=9)
        torch.cuda.empty_cache()
        forms["differentiation"] = grad_form()
        forms["differentiation_burton_miller"] = grad_bm_form()
        torch.cuda.empty_cache()
        forms["p1"] = p1_form()
        torch.cuda.empty_cache()
        forms["datagen"] = datagen_form()
        torch.cuda.empty_cache()
        forms["modal"] = modal_form()
    cpu = cpu_leg(args, v, t, ks[0])
    log(f"bench: fp64 check {fp64}")
    if forms:
        log(f"bench: forms {json.dumps(forms)}")
    out = {
        "metric": METRIC,
        "value": entries / wall,
        "unit": "entries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 (fp32)",
        "data": "synthetic (seeded perturbed icosphere, monopole Neumann data)",
        "config": {"workload": f"single large mesh: icosphere L{args.level} ({n} triangles) +5% SH perturbation, "
                               f"DP0 complex64 CBIE, k={ks}, 4 interior monopoles, GMRES tol {args.tol}",
                   "triangles": n, "basis": "DP0", "wavenumbers": ks, "gmres_restart": args.restart},
        "solves_per_s": nsolves / wall,
        "assembly_entries_per_s": 2 * n * n * len(asm) / (np.sum(asm) / 1e3),
        "gmres_iterations": its,
        "rel_err_vs_analytic": errs,
        "rel_err_vs_fp64_small_mesh": {"level": args.check_level, **fp64},
        "rel_err_vs_fp64_level6_rows": fp64_rows,
        "roofline": {"bound": "hbm", "kernel": "gemv_n_vec_kernel<4> (complex64 GEMV in GMRES)",
                     "achieved": gemv_bytes / gemv_avg_s / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": gemv_bytes / gemv_avg_s / HBM_PEAK, "traffic": _per_entry(GEMV_TRAFFIC_PER_ENTRY, n),
                     "bytes_per_launch": gemv_bytes, "mean_launch_ms": gemv_avg_s * 1e3, "launches": gemv_n,
                     "share_of_timed_region": gemv_ms / 1e3 / wall,
                     "byte_count": "8 B per matrix entry (complex64) + 16 B per row (x read, y written)"},
        "roofline_assembly": {"bound": "valu", "kernel": "dp0_cbie_kernel<6> (fused single+double layer)",
                              "achieved": achieved / 1e12, "peak": FP32_PEAK / 1e12, "unit": "TFLOP/s",
                              "frac": achieved / FP32_PEAK, "traffic": _per_entry(CBIE_TRAFFIC_PER_ENTRY, n),
                              "flop_per_launch": flop, "mean_launch_ms": reg_mean_s * 1e3,
                              "transcendentals_per_s": 4 * 36 * (n * n - npairs) / reg_mean_s,
                              "share_of_timed_region": float(np.sum(reg)) / 1e3 / wall,
                              "flop_count": "45 FLOP (single 16 + double 29) x 36 Gauss-point pairs x "
                                            "(N^2 - neighbour pairs)"},
        "roofline_matvec_standalone": mv,
        "forms": forms,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
