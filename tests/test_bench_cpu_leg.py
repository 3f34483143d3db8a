"""The benchmark's CPU leg on the host: recorded results of small DP0 CBIE
systems get their FP64 oracle errors written where the result dictionaries
expect them (per item, or the maximum over items)."""
import os
import sys

import numpy as np

import oracle
from neuralat_amd import mesh

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def _system(level=1, k=3.0):
    v, t = mesh.icosphere(level, 1.0)
    v = v.astype(np.float32)
    _, dn = mesh.sources_field([[0.1, 0.0, -0.05]], None, k)
    psi = mesh.project_dp0(v, t, lambda x, n: dn(x, n))
    S = oracle.assemble(v, t, k, "single", "DP0")
    D = oracle.assemble(v, t, k, "double", "DP0")
    M = oracle.identity(v, t, "DP0")
    phi = np.linalg.solve(0.5 * M - D, -S @ psi)
    return v, t, k, psi, phi


def test_fp64_answers_set_and_max():
    v, t, k, psi, phi = _system()
    pts = mesh.random_exterior_points(8, 1.5, 2.5, seed=1)
    p = -oracle.potential(v, t, pts, k, "single", "DP0") @ psi + oracle.potential(v, t, pts, k, "double",
                                                                                  "DP0") @ phi
    checks = bench.Fp64Checks()
    one, agg = {}, {}
    checks.add(one, v, t, k, psi, {"surface": phi}, pts, {"exterior": p * (1 + 1e-3)})
    checks.add(agg, v, t, k, psi, {"worst": phi * (1 + 2e-4)}, mode="max")
    checks.add(agg, v, t, k, psi, {"worst": phi * (1 + 5e-4)}, mode="max")
    assert one == {"surface": None, "exterior": None} and agg == {"worst": 0.0}
    bench.fp64_answers(checks)
    assert one["surface"] < 1e-12
    assert abs(one["exterior"] - 1e-3) < 1e-9
    assert abs(agg["worst"] - 5e-4) < 1e-9


def test_fp64_row_answers():
    """The level-6 row checks (kinds dp0_cbie and p1_bm) on a small mesh:
    rows scaled by 1 + 1e-4 report 1e-4, an exact solution a residual at
    rounding level."""
    v, t, k, psi, phi = _system(level=2)
    rows = np.array([0, 7, 100, 319])
    S = oracle.assemble(v, t, k, "single", "DP0")
    D = oracle.assemble(v, t, k, "double", "DP0")
    M = oracle.identity(v, t, "DP0")
    A = 0.5 * M - D
    checks = bench.Fp64Checks()
    dp0 = {}
    checks.add_rows(dp0, "dp0_cbie", v, t, k, rows, A[rows] * (1 + 1e-4), -(S @ psi)[rows], phi, psi)
    # Burton-Miller (P1), solved exactly in FP64 here
    _, dn = mesh.sources_field([[0.1, 0.0, -0.05]], None, k)
    psi1 = mesh.project_p1(v, t, lambda x, n: dn(x, n))
    M1 = oracle.identity(v, t, "P1")
    L = oracle.assemble(v, t, k, "bm_lhs", "P1", quad="approx") + 0.5 * M1
    R = oracle.assemble(v, t, k, "bm_rhs", "P1", quad="approx") - 0.5j / k * M1
    x1 = np.linalg.solve(L, R @ psi1)
    bm = {}
    rv = np.array([0, 5, 161])
    checks.add_rows(bm, "p1_bm", v, t, k, rv, L[rv], (R @ psi1)[rv] * (1 - 2e-4), x1, psi1)
    bench.fp64_answers(checks)
    assert abs(dp0["rows_rel_err_vs_fp64"] - 1e-4) < 1e-9
    assert dp0["rhs_rel_err_vs_fp64"] < 1e-12 and dp0["fp64_residual_over_rows"] < 1e-12
    assert bm["rows_rel_err_vs_fp64"] < 1e-12 and bm["fp64_residual_over_rows"] < 1e-12
    assert abs(bm["rhs_rel_err_vs_fp64"] - 2e-4) < 1e-9
