"""GPU checks of the FEM modal-analysis step against the FP64 oracle
(oracle/fem.py): block-CSR mass and stiffness matrices, the block product,
the lowest modes, and the per-triangle Neumann data of VibrationObj."""
import numpy as np
import pytest
import torch

from neuralat_amd import mesh
from oracle import fem as ofem

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def box():
    v, t = mesh.tet_box(5, 3, 4, (0.3, 0.15, 0.2), jitter=0.25, seed=7)
    return v, t


@pytest.mark.parametrize("mat", ["Ceramic", "Plastic", "Steel"])
def test_matrices_vs_oracle(box, mat):
    from neuralat_amd.modal import FEMmodel, Material, MatSet

    v, t = box
    m = Material(getattr(MatSet, mat))
    f = FEMmodel(v, t, m)
    K = f.stiffness_matrix.to_dense().cpu().numpy()
    M = f.mass_matrix.to_dense().cpu().numpy()
    v32 = v.astype(np.float32)
    Kr = ofem.assemble(v32, t, "stiffness", youngs=m.youngs, poisson=m.poison).toarray()
    Mr = ofem.assemble(v32, t, "mass", density=m.density).toarray()
    assert _rel(K, Kr) < 1e-6, mat
    assert _rel(M, Mr) < 1e-6, mat
    # the pattern holds exactly the vertex pairs that share a tetrahedron
    assert f.nblocks * 9 == np.count_nonzero(np.abs(Mr) > 0) * 3


def test_assembly_is_deterministic(box):
    from neuralat_amd.modal import FEMmodel

    v, t = box
    a = FEMmodel(v, t).blocks("stiffness").clone()
    b = FEMmodel(v, t).blocks("stiffness")
    assert torch.equal(a, b)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("m", [1, 7, 38, 70])
def test_block_product(box, m, dtype):
    from neuralat_amd.modal import FEMmodel

    v, t = box
    f = FEMmodel(v, t)
    rng = np.random.default_rng(m)
    X = rng.standard_normal((f.ndof, m)).astype(dtype)
    for which in ("mass", "stiffness"):
        Y = f.matmul(which, torch.tensor(X, device="cuda")).cpu().numpy()
        assert Y.dtype == dtype
        A = ofem.assemble(v.astype(np.float32), t, which, density=f.material.density, youngs=f.material.youngs,
                          poisson=f.material.poison)
        # the float64 blocks are the FP64 sums of the float32 vertices' records
        assert _rel(Y, A @ X.astype(np.float64)) < (1e-5 if dtype == np.float32 else 1e-6), (which, m)


def test_lowest_modes_vs_oracle():
    from neuralat_amd.modal import FEMmodel, LOBPCG_solver, Material, MatSet

    v, t = mesh.tet_box(6, 3, 2, (0.3, 0.12, 0.08), jitter=0.2, seed=2)
    m = Material(MatSet.Plastic)
    f = FEMmodel(v, t, m)
    k = 8
    w, x = LOBPCG_solver(f.stiffness_matrix, f.mass_matrix, k)
    v32 = v.astype(np.float32)
    K = ofem.assemble(v32, t, "stiffness", youngs=m.youngs, poisson=m.poison)
    M = ofem.assemble(v32, t, "mass", density=m.density)
    wr, _ = ofem.modes(K, M, k)
    assert np.all(np.abs(w - wr) <= 1e-3 * wr), (w, wr)
    # torch.lobpcg in FP32 at its default tolerance (the reference's solver):
    # eigenvalues to 1e-3, eigenvector residuals of a few per cent
    for q in range(k):
        r = K @ x[:, q] - w[q] * (M @ x[:, q])
        assert np.linalg.norm(r) <= 0.1 * w[q] * np.linalg.norm(M @ x[:, q])


@pytest.mark.parametrize("jacobi", [True, False])
def test_block_lobpcg_vs_oracle(jacobi):
    """The module's LOBPCG on the library's block products: eigenvalues to
    1e-5 of the FP64 shift-invert Lanczos modes, eigenvector residuals (in
    FP64 on the oracle's matrices) to 1e-4."""
    from neuralat_amd.modal import FEMmodel, Material, MatSet, lobpcg

    v, t = mesh.tet_box(8, 3, 3, (0.4, 0.12, 0.1), jitter=0.2, seed=4)
    m = Material(MatSet.Steel)
    f = FEMmodel(v, t, m)
    k = 10
    w, x, info = lobpcg(f, k, jacobi=jacobi)
    assert info["converged"], info
    v32 = v.astype(np.float32)
    K = ofem.assemble(v32, t, "stiffness", youngs=m.youngs, poisson=m.poison)
    M = ofem.assemble(v32, t, "mass", density=m.density)
    wr, _ = ofem.modes(K, M, k)
    assert np.all(np.abs(w - wr) <= 1e-5 * wr), (w, wr)
    for q in range(k):
        r = K @ x[:, q] - w[q] * (M @ x[:, q])
        assert np.linalg.norm(r) <= 1e-4 * wr[-1] * np.linalg.norm(M @ x[:, q]), q


def test_vibration_object_neumann():
    from neuralat_amd.modal import Material, MatSet, VibrationObj

    v, t = mesh.tet_box(6, 3, 2, (0.3, 0.12, 0.08), jitter=0.0)
    sv, st = mesh.tet_boundary(v, t)
    obj = VibrationObj(v, t, sv, st)
    obj.normalize(0.2)
    obj.modal_analysis(k=6, material=Material(MatSet.Ceramic))
    f = obj.get_frequencies()
    assert f.shape == (6,) and np.all(f > 0) and np.all(np.diff(f) >= -1e-6 * f[-1])
    assert abs(obj.get_wave_number(0) - 2 * np.pi * f[0] / 343.2) < 1e-9 * obj.get_wave_number(0)
    neu = obj.get_triangle_neumann(0)
    assert neu.shape == (len(st),)
    # the surface vertices coincide with volume vertices, so the mode at a
    # triangle is the mean of its vertices' volume mode vectors . normal
    modes = obj.modes[:, :, 0]
    want = (modes[obj.surf_triangles].mean(axis=1) * obj.surf_normals).sum(axis=1)
    assert np.allclose(neu, want)
    assert np.abs(neu).max() > 0
