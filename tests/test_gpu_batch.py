"""GPU checks of the batched form: the persistent multi-item assembly gives
exactly the per-item fused systems (same tile code and summation order, any
rows-per-task tile), the autotuner returns a candidate, and lockstep batched
GMRES (with and without graph capture) matches per-item GMRES."""
import numpy as np
import pytest
import torch

from neuralat_amd import mesh

pytestmark = pytest.mark.gpu


def _dev():
    return torch.device("cuda:0")


def _items():
    from neuralat_amd.bem.solver import BEM_Solver

    rng = np.random.default_rng(11)
    out = []
    for q, L in enumerate([3, 1, 2, 3, 2, 0]):
        v, t = mesh.icosphere(L, 1.0)
        v = mesh.anisotropic_scale(mesh.sh_perturb(v, 4, 0.05, seed=q), rng.uniform(0.5, 2.0, 3))
        k = float(np.exp(rng.uniform(np.log(5), np.log(40))))
        s = BEM_Solver(torch.tensor(v.astype(np.float32), device=_dev()), torch.tensor(t, device=_dev()),
                       basis="DP0")
        _, dn = mesh.sources_field([[0.05, -0.02, 0.01]], None, k)
        psi = torch.tensor(mesh.project_dp0(v, t, lambda x, n: dn(x, n)).astype(np.complex64), device=_dev())
        out.append((s, k, psi))
    return out


@pytest.fixture(scope="module")
def items():
    return _items()


@pytest.mark.parametrize("row_tile", [16, 32, 64])
def test_persistent_batch_equals_per_item(items, row_tile):
    from neuralat_amd.bem.batch import batch_cbie

    outs = batch_cbie([s.mesh for s, _, _ in items], [k for _, k, _ in items], [p for _, _, p in items],
                      row_tile=row_tile)
    for (s, k, psi), (lhs, rhs) in zip(items, outs):
        l1, r1 = s.cbie_system(k, psi)
        assert torch.equal(lhs, l1)
        assert torch.equal(rhs, r1)


def test_autotune_picks_a_candidate(items):
    from neuralat_amd.bem.batch import batch_autotune, batch_cbie

    best, ms = batch_autotune([s.mesh for s, _, _ in items], [k for _, k, _ in items], [p for _, _, p in items])
    assert best in (16, 32, 64) and all(v > 0 for v in ms.values())
    outs = batch_cbie([s.mesh for s, _, _ in items], [k for _, k, _ in items], [p for _, _, p in items])
    l1, r1 = items[0][0].cbie_system(items[0][1], items[0][2])
    assert torch.equal(outs[0][0], l1)


@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_gmres_matches_per_item(items, use_graph):
    from neuralat_amd.bem.batch import batch_cbie, batch_gmres
    from neuralat_amd.bem.solver import gmres

    outs = batch_cbie([s.mesh for s, _, _ in items], [k for _, k, _ in items], [p for _, _, p in items])
    xs, infos = batch_gmres([o[0] for o in outs], [o[1] for o in outs], tol=1e-6, restart=30, max_cycles=60,
                            use_graph=use_graph)
    for (lhs, rhs), x, inf in zip(outs, xs, infos):
        assert inf["converged"] and inf["rel_residual"] <= 1e-6, inf
        xr, ir = gmres(lhs, rhs, tol=1e-6, restart=100)
        assert ir["converged"]
        rel = (torch.linalg.vector_norm(x - xr) / torch.linalg.vector_norm(xr)).item()
        assert rel < 2e-5, rel


def test_batch_gmres_graph_capture_mid_cycle_equals_direct_launches():
    """Each part of a restart cycle (start, segments of 4 and then 20 Arnoldi
    steps, end) runs as direct launches on its first two uses, is captured on
    its third and replayed after.  At restart 15 (two segments) these systems
    need about 70 steps, five cycles: two direct, one captured, the rest
    replayed.
    The solutions, step counts and residuals equal those of direct launches
    bit for bit (the same kernels in the same order)."""
    from neuralat_amd.bem.batch import batch_gmres

    rng = np.random.default_rng(5)
    As, bs = [], []
    for n in (300, 180, 240):
        A = (np.eye(n) * (1.2 + 0.4j) + (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
             * (0.75 / np.sqrt(n))).astype(np.complex64)
        As.append(torch.tensor(A, device=_dev()))
        bs.append(torch.tensor((rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64),
                               device=_dev()))
    xg, ig = batch_gmres(As, bs, tol=1e-6, restart=15, max_cycles=60, use_graph=True)
    xd, idd = batch_gmres(As, bs, tol=1e-6, restart=15, max_cycles=60, use_graph=False)
    for a, b, x1, x2, i1, i2 in zip(As, bs, xg, xd, ig, idd):
        assert i1["converged"] and i1["iterations"] > 45, i1  # four cycles or more, so graphs were replayed
        assert i1["iterations"] == i2["iterations"] and i1["rel_residual"] == i2["rel_residual"], (i1, i2)
        assert torch.equal(x1, x2)
        ref = np.linalg.solve(a.cpu().numpy().astype(np.complex128), b.cpu().numpy().astype(np.complex128))
        assert np.linalg.norm(x1.cpu().numpy() - ref) / np.linalg.norm(ref) < 1e-5


def test_batch_gmres_across_fused_step_threshold():
    """Arnoldi steps whose Gram-Schmidt reads at most 65536 basis entries per
    item ((j + 1) n) run as one launch per step, deeper ones as separate
    kernels.  With n = 3000 the first 21 steps of every cycle take the fused
    launch and the rest the separate kernels; n = 40 (nmax 40) stays fused.
    Both batches reach the FP64 direct solution and the step counts of one
    GMRES per item."""
    from neuralat_amd.bem.batch import batch_gmres
    from neuralat_amd.bem.solver import gmres

    rng = np.random.default_rng(11)
    for sizes in ((3000, 2200, 1700), (40, 33)):
        As, bs = [], []
        for n in sizes:
            A = (np.eye(n) * (1.2 + 0.4j) + (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
                 * (0.6 / np.sqrt(n))).astype(np.complex64)
            As.append(torch.tensor(A, device=_dev()))
            bs.append(torch.tensor((rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64),
                                   device=_dev()))
        xs, infos = batch_gmres(As, bs, tol=1e-6, restart=60, max_cycles=20)
        for a, b, x, inf in zip(As, bs, xs, infos):
            assert inf["converged"], inf
            if a.shape[0] == 3000:
                assert inf["iterations"] > 22, inf  # crossed the threshold
            _, ir = gmres(a, b, tol=1e-6, restart=60)
            assert abs(inf["iterations"] - ir["iterations"]) <= 1, (inf, ir)
            ref = np.linalg.solve(a.cpu().numpy().astype(np.complex128), b.cpu().numpy().astype(np.complex128))
            assert np.linalg.norm(x.cpu().numpy() - ref) / np.linalg.norm(ref) < 1e-5


def test_batch_gmres_iterations_match_per_item(items):
    """Per-item Arnoldi step counts of the lockstep solve (cycles ended early
    once every item is done) agree with one GMRES per item at the same restart."""
    from neuralat_amd.bem.batch import batch_cbie, batch_gmres
    from neuralat_amd.bem.solver import gmres

    outs = batch_cbie([s.mesh for s, _, _ in items], [k for _, k, _ in items], [p for _, _, p in items])
    _, infos = batch_gmres([o[0] for o in outs], [o[1] for o in outs], tol=1e-6, restart=30, max_cycles=60)
    for (lhs, rhs), inf in zip(outs, infos):
        _, ir = gmres(lhs, rhs, tol=1e-6, restart=30)
        assert abs(inf["iterations"] - ir["iterations"]) <= max(2, 0.05 * ir["iterations"]), (inf, ir)


@pytest.mark.parametrize("dtype", ["bf16", "fp16", "fp8"])
@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_mixed_precision_matches_fp64(items, dtype, use_graph):
    """Flexible FP32 GMRES preconditioned by lockstep 16-bit or 8-bit GMRES
    reaches the FP32 residual target on every item; the solutions agree with an
    FP64 direct solve of the same systems and with the complex64 batched solve."""
    from neuralat_amd.bem.batch import batch_cbie, batch_gmres, batch_solve_mixed

    outs = batch_cbie([s.mesh for s, _, _ in items], [k for _, k, _ in items], [p for _, _, p in items])
    As, bs = [o[0] for o in outs], [o[1] for o in outs]
    xs, infos = batch_solve_mixed(As, bs, dtype=dtype, tol=1e-6, inner_tol=1e-2, outer_restart=10,
                                  inner_restart=50, max_outer_cycles=4, use_graph=use_graph)
    xr, _ = batch_gmres(As, bs, tol=1e-6, restart=30, max_cycles=60)
    for A, b, x, x32, inf in zip(As, bs, xs, xr, infos):
        assert inf["converged"] and inf["rel_residual"] <= 1e-6, inf
        assert inf["inner_iterations"] > 0 and inf["iterations"] <= 40, inf
        a64 = A.cpu().numpy().astype(np.complex128)
        x64 = np.linalg.solve(a64, b.cpu().numpy().astype(np.complex128))
        err = np.linalg.norm(x.cpu().numpy() - x64) / np.linalg.norm(x64)
        assert err < 1e-4, err
        rel = (torch.linalg.vector_norm(x - x32) / torch.linalg.vector_norm(x32)).item()
        assert rel < 1e-4, rel


def test_batch_mixed_precision_rejects_bad_dtype(items):
    from neuralat_amd.bem.batch import batch_solve_mixed

    A = torch.eye(8, dtype=torch.complex64, device=_dev())
    with pytest.raises(ValueError):
        batch_solve_mixed([A], [torch.ones(8, dtype=torch.complex64, device=_dev())], dtype="int4")


def test_batch_items_vs_fp64_oracle(items):
    """Each item of the persistent batch against the FP64 oracle directly: the
    batched system (M/2 - K, -V psi) against oracle.assemble of the item's
    mesh and wavenumber (relative Frobenius <= 1e-5), and the lockstep
    complex64 solve and the fp16 / fp8 mixed-precision solves against the
    FP64 oracle's solution of that system (<= 1e-4)."""
    import oracle
    from neuralat_amd.bem.batch import batch_cbie, batch_gmres, batch_solve_mixed

    outs = batch_cbie([s.mesh for s, _, _ in items], [k for _, k, _ in items], [p for _, _, p in items])
    As, bs = [o[0] for o in outs], [o[1] for o in outs]
    sols = {"c64": batch_gmres(As, bs, tol=1e-6, restart=30, max_cycles=60)[0]}
    for dt in ("fp16", "fp8"):
        sols[dt] = batch_solve_mixed(As, bs, dtype=dt, tol=1e-6, inner_tol=1e-2, outer_restart=10, inner_restart=50,
                                     max_outer_cycles=4)[0]
    for q, (s, k, psi) in enumerate(items):
        v = s.vertices.cpu().numpy()
        t = s.triangles.cpu().numpy()
        S = oracle.assemble(v, t, k, "single", "DP0")
        D = oracle.assemble(v, t, k, "double", "DP0")
        A64 = 0.5 * oracle.identity(v, t, "DP0") - D
        b64 = -S @ psi.cpu().numpy().astype(np.complex128)
        eA = np.linalg.norm(As[q].cpu().numpy() - A64) / np.linalg.norm(A64)
        eb = np.linalg.norm(bs[q].cpu().numpy() - b64) / np.linalg.norm(b64)
        assert eA < 1e-5 and eb < 1e-5, (q, eA, eb)
        x64 = np.linalg.solve(A64, b64)
        for name, xs in sols.items():
            e = np.linalg.norm(xs[q].cpu().numpy() - x64) / np.linalg.norm(x64)
            print(f"batch item {q} ({t.shape[0]} triangles, k = {k:.1f}) {name}: vs FP64 oracle {e:.2e}", flush=True)
            assert e < 1e-4, (q, name, e)
