"""The oracle still reproduces the committed golden fixtures (CPU only).

The fixtures (tests/golden/*.npz, written by tests/golden/gen_golden.py) are
what the GPU parity tests compare against at small sizes; this guards them
against silent drift of the oracle."""
import os

import numpy as np
import pytest

import oracle

GOLD = os.path.join(os.path.dirname(__file__), "golden")
OPS = ["single", "double", "adjointdouble", "hypersingular", "bm_lhs", "bm_rhs"]


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def test_octasphere0_p1_fixture():
    g = np.load(os.path.join(GOLD, "octasphere0_p1.npz"))
    v, t = g["vertices"], g["triangles"]
    assert _rel(oracle.identity(v, t, "P1"), g["identity"]) < 1e-14
    for k in (1, 10, 100):
        for op in OPS:
            assert _rel(oracle.assemble(v, t, float(k), op, "P1"), g[f"{op}_k{k}"]) < 1e-13, (op, k)
        assert _rel(oracle.potential(v, t, g["points"], float(k), "single", "P1") @ g["density"], g[f"spot_k{k}"]) < 1e-13


@pytest.mark.parametrize("quad", ["approx", "approx1"])
def test_octasphere1_approx_fixture(quad):
    g = np.load(os.path.join(GOLD, "octasphere1_p1_approx.npz"))
    for op in OPS:
        assert _rel(oracle.assemble(g["vertices"], g["triangles"], 10.0, op, "P1", quad=quad), g[f"{op}_{quad}"]) < 1e-13


def test_icosphere1_dp0_fixture():
    g = np.load(os.path.join(GOLD, "icosphere1_dp0.npz"))
    for op in ("single", "double", "adjointdouble", "bm_rhs"):
        assert _rel(oracle.assemble(g["vertices"], g["triangles"], 18.3, op, "DP0"), g[op]) < 1e-13
