"""GPU parity of the HIP path against the CPU oracle and the golden fixtures.

Bar: relative Frobenius error <= 1e-5 against the FP64 oracle for every
assembled operator (the plain method's FP32 parity target), with the same
r-clamp setting on both sides.  All calls go through the C ABI
(libbem_hip.so) via neuralat_amd.bem.
"""
import os

import numpy as np
import pytest
import torch

import oracle
from neuralat_amd import mesh

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
OPS = ["single", "double", "adjointdouble", "hypersingular", "bm_lhs", "bm_rhs"]
TOL = 1e-5


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _dev():
    return torch.device("cuda:0")


def _solver(v, t, basis="P1", clamp=True):
    from neuralat_amd.bem import BEM_Solver

    return BEM_Solver(torch.tensor(np.asarray(v, np.float32), device=_dev()),
                      torch.tensor(np.asarray(t, np.int32), device=_dev()), basis=basis, clamp=clamp)


def _say(*a):
    print(*a, flush=True)


def test_native_library_is_loaded():
    from neuralat_amd import _lib

    lib = _lib.load()
    assert b"gfx950" in lib.bem_version()


# ---------------------------------------------------------------------------
# golden fixtures (the reference's own check: octahedron sphere, P1, k in 1/10/100)
# ---------------------------------------------------------------------------

def test_golden_octasphere0_all_operators():
    g = np.load(os.path.join(GOLD, "octasphere0_p1.npz"))
    s = _solver(g["vertices"], g["triangles"])
    assert _rel(s.identity_matrix().cpu().numpy(), g["identity"]) < TOL
    for k in (1, 10, 100):
        for op in OPS:
            _say("golden octasphere0", op, k)
            A = s.assemble_boundary_matrix(float(k), op).cpu().numpy()
            e = _rel(A, g[f"{op}_k{k}"])
            assert e < TOL, (op, k, e)
        dens = torch.tensor(g["density"], dtype=torch.complex64, device=_dev())
        pts = torch.tensor(g["points"], device=_dev())
        for op, key, fn in (("single", "spot", s.single_potential), ("double", "dpot", s.double_potential)):
            # at k = 100 the phase k r ~ 200 is resolved to ~1e-5 in FP32, so the
            # bar is the larger of 1e-5 and 1.5x the error of the FP32 mirror of
            # the reference arithmetic on the same inputs
            mirror = oracle.potential(g["vertices"], g["triangles"], g["points"], float(k), op, "P1",
                                      precision="f32") @ g["density"]
            bar = max(TOL, 1.5 * _rel(mirror, g[f"{key}_k{k}"]))
            e = _rel(fn(float(k), dens, pts).cpu().numpy(), g[f"{key}_k{k}"])
            assert e < bar, (op, k, e, bar)


@pytest.mark.parametrize("quad", ["approx", "approx1"])
def test_golden_approx_rules(quad):
    g = np.load(os.path.join(GOLD, "octasphere1_p1_approx.npz"))
    s = _solver(g["vertices"], g["triangles"])
    for op in OPS:
        _say("golden approx", quad, op)
        A = s.assemble_boundary_matrix(10.0, op, approx=("approx_1" if quad == "approx1" else True)).cpu().numpy()
        e = _rel(A, g[f"{op}_{quad}"])
        assert e < TOL, (op, quad, e)


@pytest.mark.parametrize("op", OPS)
def test_p1_neighbour_pairs_approx_vs_oracle(op):
    """The neighbour pairs alone (parts = 2) with the 2-point line rule, which
    packs four pairs into a wavefront, grouped by kind: level 2 (320
    triangles, coincident, edge and vertex pairs in mixed groups at the kind
    boundaries) against the oracle's neighbour-pair part."""
    from neuralat_amd import _lib
    from neuralat_amd.bem.solver import _ptr, _stream

    v, t = mesh.icosphere(2, 1.0)
    v = mesh.sh_perturb(v, 4, 0.08, seed=6).astype(np.float32)
    s = _solver(v, t, "P1")
    n = len(v)
    k = 6.5
    _say("P1 neighbour pairs approx", op, s.mesh.n_neighbor_pairs)
    X = torch.zeros((n, n), dtype=torch.complex64, device=_dev())
    _lib.check(_lib.load().bem_assemble_ex(s.mesh.handle, k, _lib.OPS[op], _lib.BASES["P1"], _lib.QUADS["approx"], 1, 2,
                                           _ptr(X), n, _stream()), "bem_assemble_ex")
    R = oracle.assemble(v, t, k, op, "P1", quad="approx", parts=2, nthreads=0)
    e = _rel(X.cpu().numpy(), R)
    assert e < TOL, (op, e)


def test_golden_icosphere1_dp0():
    g = np.load(os.path.join(GOLD, "icosphere1_dp0.npz"))
    for clamp in (True, False):
        s = _solver(g["vertices"], g["triangles"], basis="DP0", clamp=clamp)
        for op in ("single", "double", "adjointdouble", "bm_rhs"):
            _say("golden icosphere1 DP0", op, clamp)
            A = s.assemble_boundary_matrix(18.3, op).cpu().numpy()
            e = _rel(A, g[op] if clamp else g[f"{op}_noclamp"])
            assert e < TOL, (op, clamp, e)
        assert _rel(s.identity_matrix().cpu().numpy(), g["identity"]) < TOL


# ---------------------------------------------------------------------------
# live oracle at larger sizes
# ---------------------------------------------------------------------------

@pytest.fixture(scope="module")
def l2_sphere():
    v, t = mesh.icosphere(2, 1.0)
    v = mesh.sh_perturb(v, 4, 0.05, seed=2)
    return v.astype(np.float32), t


@pytest.mark.parametrize("op", OPS)
def test_p1_operators_vs_oracle(l2_sphere, op):
    v, t = l2_sphere
    s = _solver(v, t, "P1")
    for k in (3.0, -12.0):
        _say("P1 L2", op, k)
        A = s.assemble_boundary_matrix(k, op).cpu().numpy()
        R = oracle.assemble(v, t, k, op, "P1")
        e = _rel(A, R)
        assert e < TOL, (op, k, e)


@pytest.mark.parametrize("op", ["single", "hypersingular", "double", "adjointdouble"])
def test_p1_many_column_tiles_vs_oracle(op):
    """Level 3 (1280 triangles): several column tiles of 256 triangles, so the
    half sweeps (row tile x column tiles from its own on; single layer and
    hypersingular: the in-place transpose-add; double layer and its adjoint:
    both blocks of each pair, the transposed ones through the scratch matrix)
    cover off-diagonal tile pairs, partial tiles and the own-tile half."""
    v, t = mesh.icosphere(3, 1.0)
    v = mesh.anisotropic_scale(mesh.sh_perturb(v, 4, 0.05, seed=4), np.array([1.3, 0.8, 1.0])).astype(np.float32)
    s = _solver(v, t, "P1")
    k = 7.0
    _say("P1 L3", op, k)
    A = s.assemble_boundary_matrix(k, op).cpu().numpy()
    R = oracle.assemble(v, t, k, op, "P1", nthreads=0)
    e = _rel(A, R)
    assert e < TOL, (op, e)


@pytest.mark.parametrize("op,quad", [("single", 0), ("hypersingular", 0), ("double", 0), ("adjointdouble", 0),
                                     ("double", 1), ("bm_lhs", 1), ("bm_rhs", 1)])
def test_p1_half_sweeps_match_ordered_pair_sweep(op, quad):
    """parts bit 4 of bem_assemble_ex evaluates every ordered triangle pair;
    the default sweeps (symmetric for the single layer and hypersingular,
    paired for the double layer and its adjoint with the 6-point rule from
    4096 vertices on) must agree with it to rounding.  Bit 8 accumulates
    directly in the output through the vertex map instead of the scratch
    matrix in the tiles' numbering: both sweeps again agree.  Level 5 (10242
    vertices, 80 column tiles); the difference is formed on the device."""
    from neuralat_amd import _lib
    from neuralat_amd.bem.solver import _ptr, _stream

    v, t = mesh.icosphere(5, 1.0)
    v = mesh.sh_perturb(v, 6, 0.05, seed=11).astype(np.float32)
    s = _solver(v, t, "P1")
    n = len(v)
    lib = _lib.load()
    out = {}
    for parts in (3, 7, 11, 15):
        _say("P1 L5 sweep", op, quad, parts)
        X = torch.empty((n, n), dtype=torch.complex64, device=_dev())
        _lib.check(lib.bem_assemble_ex(s.mesh.handle, 11.0, _lib.OPS[op], _lib.BASES["P1"], quad, 1, parts, _ptr(X),
                                       n, _stream()), "bem_assemble_ex")
        out[parts] = X
    for parts in (3, 11, 15):
        e = float(torch.linalg.matrix_norm(out[parts] - out[7]) / torch.linalg.matrix_norm(out[7]))
        assert e < 1e-6, (op, quad, parts, e)


@pytest.mark.parametrize("basis,op,quad", [("P1", "single", 0), ("P1", "double", 0), ("P1", "hypersingular", 1),
                                           ("P1", "bm_lhs", 1), ("P1", "bm_rhs", 1), ("DP0", "single", 0),
                                           ("DP0", "adjointdouble", 1)])
def test_assemble_batch_k_matches_single_calls(basis, op, quad):
    """bem_assemble_batch_k (grid z over the wavenumbers) against one
    bem_assemble per wavenumber, with ld > n and a stride larger than one
    matrix so that the offsets are exercised (level 3: 642 vertices, 1280
    triangles)."""
    from neuralat_amd import _lib
    from neuralat_amd.bem.solver import _ptr, _stream

    v, t = mesh.icosphere(3, 0.1)
    v = mesh.sh_perturb(v, 4, 0.05, seed=2).astype(np.float32)
    s = _solver(v, t, basis)
    n = len(v) if basis == "P1" else len(t)
    ld, stride = n + 3, (n + 5) * (n + 3)
    ks = np.array([3.0, 18.3, -7.5, 40.0], dtype=np.float32)
    lib = _lib.load()
    _say("batch_k", basis, op, quad)
    buf = torch.zeros(len(ks) * stride, dtype=torch.complex64, device=_dev())
    _lib.check(lib.bem_assemble_batch_k(s.mesh.handle, ks.ctypes.data, len(ks), _lib.OPS[op], _lib.BASES[basis], quad,
                                        1, _ptr(buf), ld, stride, _stream()), "bem_assemble_batch_k")
    for z, k in enumerate(ks):
        one = torch.empty((n, n), dtype=torch.complex64, device=_dev())
        _lib.check(lib.bem_assemble(s.mesh.handle, float(k), _lib.OPS[op], _lib.BASES[basis], quad, 1, _ptr(one), n,
                                    _stream()), "bem_assemble")
        got = buf[z * stride: z * stride + n * ld].reshape(n, ld)[:, :n]
        e = float(torch.linalg.matrix_norm(got - one) / torch.linalg.matrix_norm(one))
        assert e < 1e-6, (basis, op, quad, float(k), e)
        pad = buf[z * stride: (z + 1) * stride]
        assert float(pad[n * ld:].abs().max()) == 0.0  # nothing written past the matrix


def test_assemble_batch_k_level5_double_layer_vs_paired_single_calls():
    """Above 4096 vertices a single double-layer call takes the paired sweep;
    a batch of wavenumbers takes the ordered-pair sweep: the two agree to
    rounding (level 5, 10242 vertices, differences formed on the device)."""
    v, t = mesh.icosphere(5, 1.0)
    v = mesh.sh_perturb(v, 6, 0.05, seed=12).astype(np.float32)
    s = _solver(v, t, "P1")
    ks = [4.0, 15.0]
    _say("batch_k level 5 double layer")
    A = s.assemble_boundary_matrix_batch(ks, "double")
    for z, k in enumerate(ks):
        one = s.assemble_boundary_matrix(k, "double")
        e = float(torch.linalg.matrix_norm(A[z] - one) / torch.linalg.matrix_norm(one))
        del one
        assert e < 1e-6, (k, e)


def test_assemble_and_potential_batch_k_over_64_wavenumbers():
    """More wavenumbers than one launch takes (64): the host splits them; the
    potential operators of every wavenumber match single calls."""
    from neuralat_amd import _lib
    from neuralat_amd.bem.solver import _ptr, _stream

    v, t = mesh.icosphere(1, 0.1)
    v = v.astype(np.float32)
    s = _solver(v, t, "P1")
    n = len(v)
    ks = np.linspace(1.0, 60.0, 70).astype(np.float32)
    _say("batch_k 70 wavenumbers")
    A = s.assemble_boundary_matrix_batch(ks, "bm_lhs", approx=True)
    pts = torch.tensor(mesh.random_exterior_points(9, 0.2, 0.5, seed=3).astype(np.float32), device=_dev())
    Pd = s.potential_matrix_batch(ks, pts, "double")
    for z in (0, 31, 63, 64, 69):
        one = s.assemble_boundary_matrix(float(ks[z]), "bm_lhs", approx=True)
        assert float(torch.linalg.matrix_norm(A[z] - one) / torch.linalg.matrix_norm(one)) < 1e-6, z
        pd = s.potential_matrix(float(ks[z]), pts, "double")
        assert float(torch.linalg.matrix_norm(Pd[z] - pd) / torch.linalg.matrix_norm(pd)) < 1e-6, z


def test_p1_potential_point_chunks():
    """The P1 potential goes through a per-(point, triangle) scratch of at most
    512 MB per launch: 1100 points on the level-5 mesh (540 MB) take two point
    chunks; each point's entries are computed and summed the same way in any
    chunking, so the result equals two calls of 550 points bit for bit."""
    v, t = mesh.icosphere(5, 1.0)
    v = v.astype(np.float32)
    s = _solver(v, t, "P1")
    pts = torch.tensor(mesh.random_exterior_points(1100, 1.5, 3.0, seed=8).astype(np.float32), device=_dev())
    _say("P1 potential, 1100 points at level 5")
    whole = s.potential_matrix(9.0, pts, "double")
    halves = torch.cat([s.potential_matrix(9.0, pts[:550], "double"), s.potential_matrix(9.0, pts[550:], "double")])
    assert torch.equal(whole, halves)
    assert float(whole.abs().max()) > 0.0


def test_burton_miller_batch_equals_per_wavenumber_calls():
    """neumann2dirichlet_batch / boundary2potential_batch (all wavenumbers in
    one launch per kernel, one lockstep GMRES) against neumann2dirichlet and
    boundary2potential called once per wavenumber."""
    from neuralat_amd.bem import BEM_Solver

    v, t = mesh.icosphere(2, 0.1)
    v = mesh.sh_perturb(v, 4, 0.05, seed=9).astype(np.float32)
    s = BEM_Solver(torch.tensor(v, device=_dev()), torch.tensor(t, device=_dev()))
    ks = [3.0, 18.3, 41.0]
    rng = np.random.default_rng(3)
    psi = torch.tensor((rng.standard_normal((3, len(v))) + 1j * rng.standard_normal((3, len(v)))).astype(np.complex64),
                       device=_dev())
    pts = torch.tensor(mesh.random_exterior_points(20, 0.2, 0.5, seed=4).astype(np.float32), device=_dev())
    _say("Burton-Miller batch over 3 wavenumbers")
    phi, infos = s.neumann2dirichlet_batch(ks, psi, tol=1e-6)
    p = s.boundary2potential_batch(ks, psi, phi, pts)
    assert all(i["converged"] for i in infos)
    for z, k in enumerate(ks):
        d = s.neumann2dirichlet(k, psi[z], tol=1e-6)
        q = s.boundary2potential(k, psi[z], d, pts)
        assert float(torch.linalg.vector_norm(phi[z] - d) / torch.linalg.vector_norm(d)) < 1e-5, k
        assert float(torch.linalg.vector_norm(p[z] - q) / torch.linalg.vector_norm(q)) < 1e-5, k


def test_set_vertices_equals_new_handle():
    """A handle moved with bem_mesh_set_vertices assembles what a handle built
    from the new positions assembles (P1 and DP0 operators, potentials,
    normals); invalid positions are rejected and leave the handle as it was."""
    from neuralat_amd.bem import BEM_Solver

    v, t = mesh.icosphere(3, 0.1)
    v0 = mesh.sh_perturb(v, 4, 0.05, seed=5).astype(np.float32)
    rot = np.array([[0.6, -0.8, 0.0], [0.8, 0.6, 0.0], [0.0, 0.0, 1.0]])
    v1 = (mesh.anisotropic_scale(v0, np.array([1.4, 0.7, 1.1])) @ rot.T + np.array([0.3, -0.1, 0.2]))
    v1 = v1.astype(np.float32)
    tt = torch.tensor(t, device=_dev())
    pts = torch.tensor(mesh.random_exterior_points(7, 0.4, 0.6, seed=1).astype(np.float32) + np.array(
        [0.3, -0.1, 0.2], dtype=np.float32), device=_dev())
    for basis in ("P1", "DP0"):
        moved = BEM_Solver(torch.tensor(v0, device=_dev()), tt, basis=basis)
        moved.set_vertices(torch.tensor(v1, device=_dev()))
        fresh = BEM_Solver(torch.tensor(v1, device=_dev()), tt, basis=basis)
        _say("set_vertices", basis)
        assert float((moved.normals - fresh.normals).abs().max()) == 0.0
        ops = [("single", False), ("double", False)] + ([("bm_lhs", True), ("bm_rhs", True)] if basis == "P1" else [])
        for op, approx in ops:
            a = moved.assemble_boundary_matrix(11.0, op, approx=approx)
            b = fresh.assemble_boundary_matrix(11.0, op, approx=approx)
            assert float(torch.linalg.matrix_norm(a - b) / torch.linalg.matrix_norm(b)) < 1e-6, (basis, op)
        pa = moved.potential_matrix(11.0, pts, "double")
        pb = fresh.potential_matrix(11.0, pts, "double")
        assert float(torch.linalg.matrix_norm(pa - pb) / torch.linalg.matrix_norm(pb)) < 1e-6, basis
    # rejected positions: a non-finite coordinate, a collapsed triangle
    s = BEM_Solver(torch.tensor(v0, device=_dev()), tt)
    before = s.assemble_boundary_matrix(5.0, "single")
    bad = v0.copy()
    bad[3, 1] = np.nan
    with pytest.raises(ValueError):
        s.set_vertices(torch.tensor(bad, device=_dev()))
    bad = v0.copy()
    bad[t[0, 1]] = bad[t[0, 0]]
    with pytest.raises(ValueError):
        s.set_vertices(torch.tensor(bad, device=_dev()))
    assert float((s.assemble_boundary_matrix(5.0, "single") - before).abs().max()) <= 1e-6 * float(before.abs().max())


@pytest.mark.parametrize("op", ["single", "double", "adjointdouble", "bm_rhs"])
@pytest.mark.parametrize("clamp", [True, False])
def test_dp0_operators_vs_oracle(op, clamp):
    v, t = mesh.icosphere(3, 0.1)
    v = v.astype(np.float32)
    s = _solver(v, t, "DP0", clamp=clamp)
    _say("DP0 L3", op, clamp)
    A = s.assemble_boundary_matrix(18.3, op).cpu().numpy()
    R = oracle.assemble(v, t, 18.3, op, "DP0", clamp=clamp, nthreads=0)
    e = _rel(A, R)
    assert e < TOL, (op, clamp, e)


@pytest.mark.parametrize("approx", [False, True])
@pytest.mark.parametrize("clamp", [True, False])
def test_fused_cbie_system_matches_separate_operators(approx, clamp):
    from neuralat_amd.bem import matvec

    v, t = mesh.icosphere(3, 0.1)
    v = mesh.sh_perturb(v, 4, 0.05, seed=7).astype(np.float32)
    k = 18.3
    s = _solver(v, t, "DP0", clamp=clamp)
    rng = np.random.default_rng(4)
    psi = torch.tensor((rng.standard_normal(len(t)) + 1j * rng.standard_normal(len(t))).astype(np.complex64),
                       device=_dev())
    lhs, rhs = s.cbie_system(k, psi, approx=approx)
    S = s.assemble_boundary_matrix(k, "single", approx=approx)
    D = s.assemble_boundary_matrix(k, "double", approx=approx)
    M = s.identity_matrix()
    ref_lhs = (0.5 * M - D).cpu().numpy()
    ref_rhs = matvec(S, psi, alpha=-1.0).cpu().numpy()
    assert _rel(lhs.cpu().numpy(), ref_lhs) < 1e-6
    assert _rel(rhs.cpu().numpy(), ref_rhs) < 1e-6
    # and against the FP64 oracle
    quad = "approx" if approx else "full"
    S64 = oracle.assemble(v, t, k, "single", "DP0", quad=quad, clamp=clamp)
    D64 = oracle.assemble(v, t, k, "double", "DP0", quad=quad, clamp=clamp)
    M64 = oracle.identity(v, t, "DP0")
    assert _rel(lhs.cpu().numpy(), 0.5 * M64 - D64) < TOL
    assert _rel(rhs.cpu().numpy(), -S64 @ psi.cpu().numpy().astype(np.complex128)) < TOL


def test_dp0_row_blocks_match_full():
    from neuralat_amd import _lib
    from neuralat_amd.bem.solver import _ptr, _stream

    v, t = mesh.icosphere(3, 1.0)
    s = _solver(v, t, "DP0")
    full = s.assemble_boundary_matrix(10.0, "double")
    n = full.shape[0]
    blk = torch.empty((300, n), dtype=torch.complex64, device=_dev())
    _lib.check(_lib.load().bem_assemble_rows(s.mesh.handle, 10.0, 1, 0, 1, 500, 800, _ptr(blk), n, _stream()))
    torch.cuda.synchronize()
    assert torch.equal(blk, full[500:800])


def test_identity_potential_t2v_vs_oracle(l2_sphere):
    from neuralat_amd.bem import map_triangle2vertex

    v, t = l2_sphere
    for basis in ("P1", "DP0"):
        s = _solver(v, t, basis)
        assert _rel(s.identity_matrix().cpu().numpy(), oracle.identity(v, t, basis)) < TOL
        pts = mesh.random_exterior_points(40, 1.5, 3.0, seed=5).astype(np.float32)
        pt = torch.tensor(pts, device=_dev())
        for op in ("single", "double"):
            for clamp in (True, False):
                P = s.potential_matrix(7.0, pt, op, clamp=clamp).cpu().numpy()
                assert _rel(P, oracle.potential(v, t, pts, 7.0, op, basis, clamp=clamp)) < TOL, (basis, op)
    f = (np.arange(len(t)) * (0.3 - 0.1j)).astype(np.complex64)
    out = map_triangle2vertex(torch.tensor(v, device=_dev()), torch.tensor(t, device=_dev()),
                              torch.tensor(f, device=_dev())).cpu().numpy()
    assert _rel(out, oracle.triangle2vertex(len(v), t, f)) < 1e-6


def test_triangle2vertex_without_mesh_handle():
    """map_triangle2vertex builds its vertex lists on the device: equal to the
    handle-based kernel bit for bit and to the oracle's average; repeated
    calls give the same bits (the sum order does not follow the atomics);
    a vertex in no triangle gets NaN (0/0, as the reference); an index out of
    range raises instead of reading out of bounds."""
    from neuralat_amd.bem import Mesh, map_triangle2vertex
    from neuralat_amd.bem.solver import _ptr, _stream
    from neuralat_amd import _lib

    v, t = mesh.icosphere(4, 1.0)
    v = v.astype(np.float32)
    rng = np.random.default_rng(2)
    f = (rng.standard_normal(len(t)) + 1j * rng.standard_normal(len(t))).astype(np.complex64)
    vt, tt, ft = (torch.tensor(x, device=_dev()) for x in (v, t, f))
    out = map_triangle2vertex(vt, tt, ft)
    again = map_triangle2vertex(vt, tt, ft)
    assert torch.equal(out, again)
    m = Mesh(vt, tt)
    ref = torch.empty_like(out)
    _lib.check(_lib.load().bem_triangle2vertex(m.handle, _ptr(ft), _ptr(ref), _stream()))
    assert torch.equal(out, ref)
    assert _rel(out.cpu().numpy(), oracle.triangle2vertex(len(v), t, f)) < 1e-6
    # an isolated extra vertex, and a bad index
    v2 = torch.cat([vt, torch.zeros((1, 3), device=_dev())])
    out2 = map_triangle2vertex(v2, tt, ft).cpu().numpy()
    assert np.isnan(out2[-1]) and np.array_equal(out2[:-1], out.cpu().numpy())
    bad = tt.clone()
    bad[5, 1] = len(v) + 3
    with pytest.raises(ValueError):
        map_triangle2vertex(vt, bad, ft)


def test_negative_wavenumber_is_conjugate(l2_sphere):
    v, t = l2_sphere
    s = _solver(v, t, "P1")
    a = s.assemble_boundary_matrix(9.0, "single").cpu().numpy()
    b = s.assemble_boundary_matrix(-9.0, "single").cpu().numpy()
    assert _rel(b, np.conj(a)) < 1e-6


# ---------------------------------------------------------------------------
# linear algebra
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("shape", [(257, 130), (1000, 1000), (64, 4097), (3, 5)])
def test_matvec_all_ops(shape):
    from neuralat_amd.bem import matvec

    rng = np.random.default_rng(1)
    r, c = shape
    A = (rng.standard_normal((r, c)) + 1j * rng.standard_normal((r, c))).astype(np.complex64)
    At = torch.tensor(A, device=_dev())
    for trans, Aop in ((0, A), (1, A.T), (2, A.conj().T)):
        x = (rng.standard_normal(Aop.shape[1]) + 1j * rng.standard_normal(Aop.shape[1])).astype(np.complex64)
        y0 = (rng.standard_normal(Aop.shape[0]) + 1j * rng.standard_normal(Aop.shape[0])).astype(np.complex64)
        ref = (0.5 - 2j) * (Aop.astype(np.complex128) @ x) + (1 + 1j) * y0
        out = torch.tensor(y0, device=_dev())
        matvec(At, torch.tensor(x, device=_dev()), trans=trans, alpha=0.5 - 2j, beta=1 + 1j, out=out)
        assert _rel(out.cpu().numpy(), ref) < 1e-5, (shape, trans)


def test_gmres_and_bicgstab_vs_direct():
    from neuralat_amd.bem import bicgstab, gmres

    rng = np.random.default_rng(2)
    n = 700
    A = (np.eye(n) * 2 + (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(n) * 0.6)
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    xr = np.linalg.solve(A, b)
    At = torch.tensor(A.astype(np.complex64), device=_dev())
    bt = torch.tensor(b.astype(np.complex64), device=_dev())
    for restart in (10, 200):
        x, info = gmres(At, bt, tol=1e-6, restart=restart, maxiter=500)
        assert info["converged"] and info["rel_residual"] <= 1e-6, info
        assert _rel(x.cpu().numpy(), xr) < 1e-5
    x, info = bicgstab(At, bt)
    assert info["converged"], info
    assert _rel(x.cpu().numpy(), xr) < 2e-5
    # silent non-convergence, as BiCGSTAB.py
    x1, info1 = bicgstab(At, bt, nsteps=1)
    assert info1["iterations"] == 1 and not info1["converged"]
    # GMRES reports non-convergence
    x2, info2 = gmres(At, bt, tol=1e-12, restart=5, maxiter=5)
    assert not info2["converged"] and info2["iterations"] == 5
    # the timed entry point runs the same solve (bit for bit) and counts one
    # product per Arnoldi step plus one explicit residual per restart cycle and
    # a final one (a cycle may stop on its implicit residual and be followed by
    # another when the explicit one is still above the tolerance)
    x3, info3 = gmres(At, bt, tol=1e-6, restart=10, maxiter=500, time_products=True)
    x4, info4 = gmres(At, bt, tol=1e-6, restart=10, maxiter=500)
    assert torch.equal(x3, x4) and info3["iterations"] == info4["iterations"]
    its3 = info3["iterations"]
    assert its3 + -(-its3 // 10) + 1 <= info3["gemv_count"] <= 2 * its3 + 1
    assert info3["gemv_ms"] > 0.0
    # transposed solves (the adjoint of the differentiation form) against numpy
    for trans, Aop in ((1, A.T), (2, A.conj().T)):
        from neuralat_amd import _lib
        from neuralat_amd.bem.solver import _ptr, _stream
        import ctypes

        lib = _lib.load()
        xt = torch.zeros(n, dtype=torch.complex64, device=_dev())
        wk = torch.empty(lib.bem_gmres_work_size(n, 60), dtype=torch.uint8, device=_dev())
        inf = _lib.SolveInfo()
        _lib.check(lib.bem_gmres_ex(_ptr(At), n, n, trans, _ptr(bt), _ptr(xt), 1e-6, 60, 500, _ptr(wk),
                                    ctypes.byref(inf), _stream()))
        assert _rel(xt.cpu().numpy(), np.linalg.solve(Aop, b)) < 1e-5, trans


def test_solve_linear_equation_matches_reference_semantics():
    from neuralat_amd.bem import solve_linear_equation

    rng = np.random.default_rng(3)
    n = 300
    A = (np.eye(n) * 3 + (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(n) * 0.4)
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    x_or, steps, conv = oracle.bicgstab(lambda z: A @ z, b)
    At = torch.tensor(A.astype(np.complex64), device=_dev())
    bt = torch.tensor(b.astype(np.complex64), device=_dev())
    x = solve_linear_equation(At, bt).cpu().numpy()
    assert _rel(x, x_or) < 2e-5
    xc = solve_linear_equation(lambda z: At @ z, bt).cpu().numpy()
    assert _rel(xc, x_or) < 2e-5


# ---------------------------------------------------------------------------
# end to end: CBIE against the oracle solution and the analytic monopole
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("basis", ["DP0", "P1"])
def test_cbie_monopole_end_to_end(basis):
    v, t = mesh.icosphere(3, 0.1)
    k = 18.3
    val, dn = mesh.sources_field([[0.01, -0.005, 0.002]], None, k)
    proj = mesh.project_dp0 if basis == "DP0" else mesh.project_p1
    psi = proj(v, t, lambda x, n: dn(x, n))
    s = _solver(v, t, basis, clamp=False)
    psit = torch.tensor(psi.astype(np.complex64), device=_dev())
    phi = s.CBIE(k, psit)
    assert s.last_solve_info["converged"]
    pts = mesh.random_exterior_points(64, 0.2, 0.5, seed=0)
    p = s.boundary2potential(k, psit, phi, torch.tensor(pts, dtype=torch.float32, device=_dev())).cpu().numpy()
    err_an = _rel(p, val(pts))
    # FP64 oracle solution of the same discrete system
    v32 = v.astype(np.float32)
    S = oracle.assemble(v32, t, k, "single", basis, clamp=False)
    D = oracle.assemble(v32, t, k, "double", basis, clamp=False)
    M = oracle.identity(v32, t, basis)
    phi_or = np.linalg.solve(0.5 * M - D, -S @ psi)
    p_or = (-oracle.potential(v32, t, pts, k, "single", basis, clamp=False) @ psi
            + oracle.potential(v32, t, pts, k, "double", basis, clamp=False) @ phi_or)
    err_fp64 = _rel(p, p_or)
    err_surf = _rel(phi.cpu().numpy(), phi_or)
    _say(f"CBIE {basis}: vs analytic {err_an:.2e}, vs FP64 oracle {err_fp64:.2e} (surface {err_surf:.2e})")
    assert err_fp64 < 1e-4 and err_surf < 1e-4
    assert err_an < 1e-3


# ---------------------------------------------------------------------------
# size-independent properties at a large size, and error paths
# ---------------------------------------------------------------------------

def test_gauss_law_large_mesh():
    """Row sums of the k=0 DLP are -|T_i|/2 on a closed surface (exact on a
    polyhedron up to quadrature error), checked at level 5 (20480 panels)."""
    from neuralat_amd.bem import matvec

    v, t = mesh.icosphere(5, 1.0)
    s = _solver(v, t, "DP0", clamp=False)
    K = s.assemble_boundary_matrix(0.0, "double")
    ones = torch.ones(K.shape[0], dtype=torch.complex64, device=_dev())
    rs = matvec(K, ones).cpu().numpy()
    area = 0.5 * np.linalg.norm(np.cross(v[t[:, 1]] - v[t[:, 0]], v[t[:, 2]] - v[t[:, 0]]), axis=1)
    assert np.abs(rs.real / area + 0.5).max() < 1e-3
    S = s.assemble_boundary_matrix(5.0, "single")
    d = (S - S.T).abs().max().item() / S.abs().max().item()
    assert d < 1e-3
    del K, S


def test_bad_meshes_raise():
    from neuralat_amd.bem import BEM_Solver

    v = torch.tensor([[0, 0, 0], [1, 0, 0], [0, 1, 0], [2, 2, 2]], dtype=torch.float32, device=_dev())
    with pytest.raises(ValueError):
        BEM_Solver(v, torch.tensor([[0, 1, 9]], dtype=torch.int32, device=_dev()))
    with pytest.raises(ValueError):
        BEM_Solver(v, torch.tensor([[0, 1, 1]], dtype=torch.int32, device=_dev()))
    vz = torch.tensor([[0, 0, 0], [1, 0, 0], [2, 0, 0]], dtype=torch.float32, device=_dev())
    with pytest.raises(ValueError):
        BEM_Solver(vz, torch.tensor([[0, 1, 2]], dtype=torch.int32, device=_dev()))
    with pytest.raises(AssertionError):
        BEM_Solver(v.double(), torch.tensor([[0, 1, 2]], dtype=torch.int32, device=_dev()))
    s = BEM_Solver(v, torch.tensor([[0, 1, 2], [0, 2, 3]], dtype=torch.int32, device=_dev()), basis="DP0")
    with pytest.raises(ValueError):
        s.assemble_boundary_matrix(1.0, "hypersingular")
    with pytest.raises(ValueError):
        s.assemble_boundary_matrix(1.0, "nonsense")


@pytest.mark.parametrize("basis", ["P1", "DP0"])
def test_potential_more_points_than_one_launch(basis):
    """65535 x 16 = 1048560 points fill one launch's grid; more are split
    into several launches.  Entries of points on either side of the split
    (and the last point) equal the same points evaluated in a small call,
    bit for bit: each (point, triangle) term is formed the same way."""
    v, t = mesh.icosphere(0, 1.0)
    s = _solver(v.astype(np.float32), t, basis)
    npts = 1048560 + 1000
    rng = np.random.default_rng(17)
    d = rng.standard_normal((npts, 3))
    pts = (d / np.linalg.norm(d, axis=1, keepdims=True) * rng.uniform(1.5, 3.0, (npts, 1))).astype(np.float32)
    pt = torch.tensor(pts, device=_dev())
    _say("potential beyond one launch", basis, npts)
    idx = torch.tensor([0, 1, 1048559, 1048560, 1048561, npts - 1], device=_dev())
    for op in ("single", "double"):
        P = s.potential_matrix(5.0, pt, op)
        Q = s.potential_matrix(5.0, pt[idx], op)
        assert P.shape[0] == npts
        assert torch.equal(P[idx], Q), (basis, op)
        del P


def test_single_triangle_mesh():
    v = np.array([[0, 0, 0], [0.1, 0, 0], [0.02, 0.09, 0]], dtype=np.float32)
    t = np.array([[0, 1, 2]], dtype=np.int32)
    for basis in ("P1", "DP0"):
        s = _solver(v, t, basis, clamp=False)
        A = s.assemble_boundary_matrix(4.0, "single").cpu().numpy()
        R = oracle.assemble(v, t, 4.0, "single", basis, clamp=False)
        assert _rel(A, R) < TOL


@pytest.mark.parametrize("n,restart", [(300, 40), (300, 300), (5000, 60)])
def test_gmres_windowed_residual_reads(n, restart):
    """Small systems read the implicit residuals every few steps (the fused
    one-workgroup step); the iterations, convergence and solution equal the
    per-step reads of the timed path, also across restarts; larger systems
    (5000 unknowns) take the multi-kernel step."""
    from neuralat_amd.bem.solver import gmres

    rng = np.random.default_rng(n + restart)
    A = (np.eye(n) * (1.5 + 0.5j) + (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
         * (0.9 / np.sqrt(n))).astype(np.complex64)
    b = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    At, bt = torch.tensor(A, device="cuda"), torch.tensor(b, device="cuda")
    x1, i1 = gmres(At, bt, tol=1e-6, restart=restart, maxiter=2000)
    x2, i2 = gmres(At, bt, tol=1e-6, restart=restart, maxiter=2000, time_products=True)
    assert i1["converged"] and i2["converged"]
    assert i1["iterations"] == i2["iterations"], (i1, i2)
    assert torch.allclose(x1, x2, rtol=0, atol=1e-6 * float(torch.linalg.vector_norm(x2)))
    ref = np.linalg.solve(A.astype(np.complex128), b.astype(np.complex128))
    assert np.linalg.norm(x1.cpu().numpy() - ref) / np.linalg.norm(ref) < 1e-5


def test_gmres_small_maxiter_and_exact_convergence():
    """maxiter inside a read window stops at maxiter (not the window's end);
    an exactly solvable Krylov space (A = I + rank 2) converges in its first
    steps, the breakdown steps after it dropped."""
    from neuralat_amd.bem.solver import gmres

    rng = np.random.default_rng(3)
    n = 200
    A = (np.eye(n) + 0.9 * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(n)
         ).astype(np.complex64)
    b = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    At, bt = torch.tensor(A, device="cuda"), torch.tensor(b, device="cuda")
    _, info = gmres(At, bt, tol=1e-12, restart=50, maxiter=11)
    assert not info["converged"] and info["iterations"] == 11, info
    u = rng.standard_normal((n, 2)) + 1j * rng.standard_normal((n, 2))
    B = (np.eye(n) + u @ u.conj().T / n).astype(np.complex64)
    x, info = gmres(torch.tensor(B, device="cuda"), bt, tol=1e-6, restart=50, maxiter=100)
    assert info["converged"] and info["iterations"] <= 4, info
    ref = np.linalg.solve(B.astype(np.complex128), b.astype(np.complex128))
    assert np.linalg.norm(x.cpu().numpy() - ref) / np.linalg.norm(ref) < 1e-5
