"""The FP64 reference of the level-4 gradient check (tests/fp64_grad.py)
against whole-pipeline central differences of the FP64 oracle (CPU only): the
linearised derivative along single-vertex directions equals the difference
quotient of two full FP64 solves to its truncation error."""
import numpy as np

import fp64_grad
from neuralat_amd import mesh


def test_single_vertex_derivative_matches_whole_pipeline_differences():
    v, t = mesh.icosphere(1, 1.0)
    v = mesh.sh_perturb(v, 4, 0.05, seed=1)
    pts = mesh.random_exterior_points(8, 1.5, 3.0, seed=2)
    src = np.array([0.1, -0.05, 0.08])
    k = 6.0
    rng = np.random.default_rng(3)
    dirs = [(int(w), rng.standard_normal(3)) for w in rng.choice(len(v), 4, replace=False)]
    lin = fp64_grad.directional_derivatives(v, t, k, src, pts, dirs)
    h = 1e-6
    fd = []
    for w, d in dirs:
        vp, vm = v.copy(), v.copy()
        vp[w] += h * d
        vm[w] -= h * d
        fd.append((fp64_grad.loss(vp, t, k, src, pts) - fp64_grad.loss(vm, t, k, src, pts)) / (2 * h))
    fd = np.array(fd)
    assert np.all(np.abs(lin - fd) <= 1e-6 * np.abs(fd)), (lin, fd)
