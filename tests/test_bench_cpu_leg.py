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
