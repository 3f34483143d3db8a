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

The ``roofline`` object is for the dominant kernel, the fused regular-pair
DP0 kernel that evaluates the double layer (stored as -K) and the single
layer (contracted with psi): algorithmic FLOP per launch = (29 + 16) FLOP x
36 Gauss-point pairs x (N^2 - neighbour pairs), divided by its mean duration
from HIP events recorded on the stream it runs on.

``cpu_baseline``: the CPU oracle's FP32 mirror of the reference arithmetic
(OpenMP, all cores of this process's affinity) timed on a row block of the same
operators; the reference's own CPU path (bempp-cl/numba) is not installed.
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
DLP_FLOP_PER_EVAL = 29  # algorithmic count per Gauss-point pair (SLP: 16)
SLP_FLOP_PER_EVAL = 16


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=2)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--level", type=int, default=6)
    p.add_argument("--ks", type=str, default="10,30,60")
    p.add_argument("--tol", type=float, default=1e-6)
    p.add_argument("--restart", type=int, default=200)
    p.add_argument("--maxiter", type=int, default=3000)
    p.add_argument("--cpu-seconds", type=float, default=15.0)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--check-level", type=int, default=3, help="mesh level of the FP64 oracle check (0 = skip)")
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
        phi, info = gmres(self.lhs, self.rhs, tol=tol, restart=restart, maxiter=maxiter)
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


def fp64_check(level, ks):
    """Same pipeline on a small mesh against the FP64 oracle solution."""
    import oracle
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
        S = oracle.assemble(v32, t, k, "single", "DP0")
        D = oracle.assemble(v32, t, k, "double", "DP0")
        M = oracle.identity(v32, t, "DP0")
        phi64 = np.linalg.solve(0.5 * M - D, -S @ psi)
        p64 = -oracle.potential(v32, t, pts, k, "single", "DP0") @ psi + oracle.potential(v32, t, pts, k, "double",
                                                                                           "DP0") @ phi64
        out[str(k)] = {
            "surface_rel_err": float(np.linalg.norm(phi.cpu().numpy() - phi64) / np.linalg.norm(phi64)),
            "exterior_rel_err": float(np.linalg.norm(p.cpu().numpy() - p64) / np.linalg.norm(p64)),
        }
    del s
    return out


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
    log(f"bench: building level-{args.level} workload")
    v, t, src, pts = make_workload(args.level)
    n = t.shape[0]
    fp64 = {}
    if args.check_level > 0:
        log(f"bench: FP64 oracle check on level {args.check_level}")
        fp64 = fp64_check(args.check_level, ks)
        log(f"bench: fp64 check {fp64}")
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

    def step(record):
        res = []
        for k in ks:
            phi, p, info = s.solve(k, psis[k], pts_t, args.tol, args.restart, args.maxiter)
            reg_ms, asm_ms = s.timings()
            res.append((k, p, info, reg_ms, asm_ms))
            if record:
                log(f"bench: k={k} fused regular kernel {reg_ms:.1f} ms, assembly {asm_ms:.1f} ms, "
                    f"GMRES its {info['iterations']} res {info['rel_residual']:.2e}")
        return res

    for w in range(args.warmup):
        log(f"bench: warmup step {w}")
        step(True)
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
    cpu = None
    if not args.no_cpu:
        log("bench: CPU baseline")
        cpu = cpu_baseline(v, t, ks[0], args.cpu_seconds)
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
        "roofline": {"bound": "valu", "kernel": "dp0_cbie_kernel<6> (fused single+double layer)", "achieved": achieved / 1e12,
                     "peak": FP32_PEAK / 1e12, "unit": "TFLOP/s", "frac": achieved / FP32_PEAK, "traffic": None,
                     "flop_per_launch": flop, "mean_launch_ms": reg_mean_s * 1e3,
                     "transcendentals_per_s": 4 * 36 * (n * n - npairs) / reg_mean_s,
                     "flop_count": "45 FLOP (single 16 + double 29) x 36 Gauss-point pairs x (N^2 - neighbour pairs)"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
