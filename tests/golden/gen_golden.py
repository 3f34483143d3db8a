"""Writes the golden fixtures in this folder from the FP64 CPU oracle.

The reference ships no golden vectors and cannot be run here (CUDA + bempp;
see the oracle's header), so the fixtures are oracle outputs, written after
the oracle itself is pinned by tests/test_oracle_*.py.  Inputs mirror the
reference's own checks: the refined-octahedron sphere at level 0 with a P1
space at k in {1, 10, 100} for all six operators, the identity and the
potentials at (0,0,2), (0,2,0), (2,0,0) (scripts/check_with_bempp.py:18-220),
plus DP0 and approximate-rule cases on icospheres.

    python tests/golden/gen_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402
from neuralat_amd import mesh  # noqa: E402

OPS = ["single", "double", "adjointdouble", "hypersingular", "bm_lhs", "bm_rhs"]


def main():
    # 1. octahedron sphere, P1, the reference's matrix check
    v, t = mesh.octasphere(0, 1.0)
    v32 = v.astype(np.float32)
    out = {"vertices": v32, "triangles": t, "ks": np.array([1.0, 10.0, 100.0])}
    out["identity"] = oracle.identity(v32, t, "P1")
    for k in (1.0, 10.0, 100.0):
        for op in OPS:
            out[f"{op}_k{int(k)}"] = oracle.assemble(v32, t, k, op, "P1")
    pts = np.array([[0.0, 0.0, 2.0], [0.0, 2.0, 0.0], [2.0, 0.0, 0.0]], dtype=np.float32)
    rng = np.random.default_rng(0)
    dens = rng.random(len(v)).astype(np.complex128)
    out["points"] = pts
    out["density"] = dens
    for k in (1.0, 10.0, 100.0):
        out[f"spot_k{int(k)}"] = oracle.potential(v32, t, pts, k, "single", "P1") @ dens
        out[f"dpot_k{int(k)}"] = oracle.potential(v32, t, pts, k, "double", "P1") @ dens
    np.savez_compressed(os.path.join(HERE, "octasphere0_p1.npz"), **out)

    # 2. octahedron sphere level 1, P1, approximate rules
    v, t = mesh.octasphere(1, 1.0)
    v32 = v.astype(np.float32)
    out = {"vertices": v32, "triangles": t}
    for quad in ("approx", "approx1"):
        for op in OPS:
            out[f"{op}_{quad}"] = oracle.assemble(v32, t, 10.0, op, "P1", quad=quad)
    np.savez_compressed(os.path.join(HERE, "octasphere1_p1_approx.npz"), **out)

    # 3. icosphere level 1 scaled to 0.1 m, DP0 at k = 18.3 (with and without clamp)
    v, t = mesh.icosphere(1, 0.1)
    v32 = v.astype(np.float32)
    out = {"vertices": v32, "triangles": t, "k": np.array(18.3)}
    for op in ("single", "double", "adjointdouble", "bm_rhs"):
        out[f"{op}"] = oracle.assemble(v32, t, 18.3, op, "DP0")
        out[f"{op}_noclamp"] = oracle.assemble(v32, t, 18.3, op, "DP0", clamp=False)
    out["identity"] = oracle.identity(v32, t, "DP0")
    np.savez_compressed(os.path.join(HERE, "icosphere1_dp0.npz"), **out)
    print("wrote fixtures to", HERE)


if __name__ == "__main__":
    main()
