"""Batched data generation for the neural acoustic-transfer datasets.

Mirrors the static-object part of the reference's scene sampler and dataset
writer (src/scene.py:114-291, experiments/neuPAT/generate.py:13-39): each
sample places the objects (resize, rotation about a pivot, move, position),
merges them into one mesh with per-vertex Neumann data from the objects'
vibration masks, draws a frequency log-uniformly in [freq_min, freq_max]
(k = 2 pi f / 343.2), solves the P1 Burton-Miller system of
``BEM_Solver.neumann2dirichlet`` and evaluates |p| at target points sampled in
a spherical shell around the target box.  The samples of a batch share the
scene's triangles, so they form one geometry set: each operator kind of
every sample is assembled in one launch per kernel, the systems are solved
together by the lockstep batched GMRES (one graph-captured restart cycle for
all of them), and the pressures of all samples come from one launch per
kernel as well (the reference solves sample by sample, src/scene.py:270-291).

Output layout, as generate.py: x [S, T, 3 + n_rot + n_move + 2] float32 =
(target factors, rotation factors, move factors, resize factor, frequency
factor), y [S, T, 1] float32 = |p|.  Meshes are given as arrays (file loading
and trajectories/object lists of the reference are outside the BEM path).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import torch

from .bem.batch import batch_gmres
from .bem.solver import BEM_Solver, GeometrySet, Mesh, boundary2potential_samples, burton_miller_samples

SPEED_OF_SOUND = 343.2  # src/scene.py:275


@dataclass
class SceneObject:
    """One static object (src/scene.py:114-191).  ``resize``: per-axis scale
    weights (a zero entry keeps that axis), ``rot_axis``: scipy Euler axis
    ("x", "y", "z"), ``rot_pos``: pivot (scaled with the object),
    ``rot_max_deg``, ``move``: displacement at factor 1, ``position``: offset,
    ``vibration``: "x", "-y", ... marks the vibrating half (Neumann 1)."""

    vertices: np.ndarray
    triangles: np.ndarray
    resize: tuple | None = None
    rot_axis: str | None = None
    rot_pos: tuple | None = None
    rot_max_deg: float | None = None
    move: tuple | None = None
    position: tuple = (0.0, 0.0, 0.0)
    vibration: str | None = None
    neumann: np.ndarray = field(init=False)

    def __post_init__(self):
        self.vertices = np.asarray(self.vertices, dtype=np.float64)
        self.triangles = np.asarray(self.triangles, dtype=np.int32)
        self.neumann = np.zeros(len(self.vertices), dtype=np.complex64)
        if self.vibration is not None:
            axis = {"x": 0, "y": 1, "z": 2}[self.vibration.replace("-", "")]
            coord = self.vertices[:, axis]
            self.neumann[(coord < 0) if self.vibration.startswith("-") else (coord > 0)] = 1.0

    def place(self, resize_factor, rot_factor, move_factor, moving_scene=True):
        """Vertices for the given factors: resize, rotation about the scaled
        pivot, then move and position -- the position offset is applied by the
        move step, which the reference runs only in scenes that have a moving
        object (src/scene.py:186-190,243-249)."""
        from scipy.spatial.transform import Rotation

        base = np.zeros(3) if self.resize is None else np.asarray(self.resize, dtype=np.float64)
        scale = base * resize_factor
        scale[scale == 0] = 1.0
        v = self.vertices * scale
        if self.rot_axis is not None:
            pivot = np.asarray(self.rot_pos, dtype=np.float64) * scale
            v = Rotation.from_euler(self.rot_axis, rot_factor * self.rot_max_deg, degrees=True).apply(v - pivot) + pivot
        if moving_scene:
            if self.move is not None:
                v = v + np.asarray(self.move, dtype=np.float64) * move_factor
            v = v + np.asarray(self.position, dtype=np.float64)
        return v

    def place_batch(self, resize_factors, rot_factors, move_factors, moving_scene=True):
        """``place`` for S samples at once (factor arrays of length S):
        vertices [S, V, 3]."""
        from scipy.spatial.transform import Rotation

        base = np.zeros(3) if self.resize is None else np.asarray(self.resize, dtype=np.float64)
        scale = base[None, :] * np.asarray(resize_factors, dtype=np.float64)[:, None]
        scale[scale == 0] = 1.0
        v = self.vertices[None] * scale[:, None, :]
        if self.rot_axis is not None:
            pivot = (np.asarray(self.rot_pos, dtype=np.float64)[None, :] * scale)[:, None, :]
            ang = np.asarray(rot_factors, dtype=np.float64) * self.rot_max_deg
            R = Rotation.from_euler(self.rot_axis, ang, degrees=True).as_matrix()  # [S, 3, 3]
            v = np.matmul(v - pivot, np.ascontiguousarray(R.transpose(0, 2, 1))) + pivot
        if moving_scene:
            if self.move is not None:
                v = v + np.asarray(self.move, dtype=np.float64) * np.asarray(move_factors, dtype=np.float64)[:, None,
                                                                                                           None]
            v = v + np.asarray(self.position, dtype=np.float64)
        return v


class SceneSampler:
    """Samples scenes and solves them in batches (src/scene.py:193-291)."""

    def __init__(self, objects, trg_sample_num, freq_min, freq_max, trg_pos_min, trg_pos_max, max_resize=2.0):
        self.objects = list(objects)
        self.rot_num = sum(o.rot_axis is not None for o in self.objects)
        self.move_num = sum(o.move is not None for o in self.objects)
        self.trg_sample_num = int(trg_sample_num)
        self.freq_min_log, self.freq_max_log = np.log10(freq_min), np.log10(freq_max)
        lo, hi = np.asarray(trg_pos_min, dtype=np.float64), np.asarray(trg_pos_max, dtype=np.float64)
        self.bbox_size = float((hi - lo).max())
        self.bbox_center = (hi + lo) / 2
        self.max_resize = float(max_resize)

    @property
    def feature_dim(self):
        return 3 + self.rot_num + self.move_num + 2

    def sample_factors(self, rng):
        """Uniform [0, 1) factors of one sample, in the order the reference draws them."""
        return {"rot": rng.random(self.rot_num), "move": rng.random(self.move_num), "resize": rng.random(),
                "freq": rng.random(), "trg": rng.random((self.trg_sample_num, 3))}

    def build(self, f):
        """(vertices, triangles, neumann, k, target points) of one sample (src/scene.py:218-291)."""
        verts, tris, neu = [], [], []
        ri = mi = 0
        nv = 0
        for o in self.objects:
            rf = f["rot"][ri] if o.rot_axis is not None else 0.0
            mf = f["move"][mi] if o.move is not None else 0.0
            v = o.place(f["resize"] * (self.max_resize - 1) + 1, rf, mf, self.move_num > 0)
            ri += o.rot_axis is not None and ri < self.rot_num - 1
            mi += o.move is not None and mi < self.move_num - 1
            verts.append(v)
            tris.append(o.triangles + nv)
            neu.append(o.neumann)
            nv += len(v)
        freq = 10 ** (f["freq"] * (self.freq_max_log - self.freq_min_log) + self.freq_min_log)
        k = 2 * np.pi * freq / SPEED_OF_SOUND
        u = f["trg"]
        rs = (u[:, 0] + 1) * self.bbox_size
        theta = u[:, 1] * 2 * np.pi - np.pi
        phi = u[:, 2] * np.pi
        pts = np.stack([rs * np.sin(phi) * np.cos(theta), rs * np.sin(phi) * np.sin(theta), rs * np.cos(phi)], -1)
        return (np.concatenate(verts), np.concatenate(tris), np.concatenate(neu), float(k),
                pts + self.bbox_center)

    def build_batch(self, facs):
        """``build`` for a list of samples at once: (vertices [S, V, 3],
        triangles, neumann [S, V], k [S], target points [S, T, 3]); the
        samples share the triangles."""
        S = len(facs)
        resize = np.array([f["resize"] for f in facs]) * (self.max_resize - 1) + 1
        verts, tris, neu = [], [], []
        ri = mi = 0
        nv = 0
        for o in self.objects:
            rf = np.array([f["rot"][ri] for f in facs]) if o.rot_axis is not None else np.zeros(S)
            mf = np.array([f["move"][mi] for f in facs]) if o.move is not None else np.zeros(S)
            v = o.place_batch(resize, rf, mf, self.move_num > 0)
            ri += o.rot_axis is not None and ri < self.rot_num - 1
            mi += o.move is not None and mi < self.move_num - 1
            verts.append(v)
            tris.append(o.triangles + nv)
            neu.append(o.neumann)
            nv += v.shape[1]
        fr = np.array([f["freq"] for f in facs])
        freq = 10 ** (fr * (self.freq_max_log - self.freq_min_log) + self.freq_min_log)
        k = 2 * np.pi * freq / SPEED_OF_SOUND
        u = np.stack([f["trg"] for f in facs])
        rs = (u[..., 0] + 1) * self.bbox_size
        theta = u[..., 1] * 2 * np.pi - np.pi
        phi = u[..., 2] * np.pi
        pts = np.stack([rs * np.sin(phi) * np.cos(theta), rs * np.sin(phi) * np.sin(theta), rs * np.cos(phi)], -1)
        return (np.concatenate(verts, axis=1), np.concatenate(tris), np.broadcast_to(np.concatenate(neu), (S, nv)),
                k, pts + self.bbox_center)

    def features_batch(self, facs):
        """``features`` of a list of samples: [S, T, D]."""
        x = np.zeros((len(facs), self.trg_sample_num, self.feature_dim), dtype=np.float32)
        x[:, :, :3] = np.stack([f["trg"] for f in facs])
        if self.rot_num:
            x[:, :, 3:3 + self.rot_num] = np.stack([f["rot"] for f in facs])[:, None, :]
        if self.move_num:
            x[:, :, 3 + self.rot_num:3 + self.rot_num + self.move_num] = np.stack([f["move"] for f in facs])[:, None, :]
        x[:, :, -2] = np.array([f["resize"] for f in facs])[:, None]
        x[:, :, -1] = np.array([f["freq"] for f in facs])[:, None]
        return x

    def features(self, f):
        x = np.zeros((self.trg_sample_num, self.feature_dim), dtype=np.float32)
        x[:, :3] = f["trg"]
        x[:, 3:3 + self.rot_num] = f["rot"]
        x[:, 3 + self.rot_num:3 + self.rot_num + self.move_num] = f["move"]
        x[:, -2] = f["resize"]
        x[:, -1] = f["freq"]
        return x

    def batch_size(self, max_batch=64, max_batch_bytes=64 << 30):
        """Samples per geometry set: up to ``max_batch``, as many as keep the
        batch's dense systems (n^2 complex64 entries each, n the scene's vertex
        count) within ``max_batch_bytes``.  One lockstep GMRES call per batch:
        on small scenes the launches and host round trips of a call, not the
        arithmetic, set its time, so larger batches are faster."""
        n = sum(len(o.vertices) for o in self.objects)
        return int(max(1, min(max_batch, max_batch_bytes // (8 * n * n))))

    def generate(self, n_samples, seed=0, batch=None, device="cuda", tol=1e-6, restart=100, max_cycles=20):
        """x [S, T, D], y [S, T, 1] (generate.py:13-39), ``batch`` samples per
        geometry set and lockstep GMRES call (default ``batch_size()``).
        Returns (x, y, infos)."""
        rng = np.random.default_rng(seed)
        dev = torch.device(device)
        S, T = int(n_samples), self.trg_sample_num
        batch = self.batch_size() if batch is None else int(batch)
        if batch < 1:
            raise ValueError("batch must be at least 1")
        x = torch.zeros((S, T, self.feature_dim), dtype=torch.float32)
        y = torch.zeros((S, T, 1), dtype=torch.float32)
        infos = []
        mesh = None  # the scene's triangles (every sample has the same ones)
        for s0 in range(0, S, batch):
            facs = [self.sample_factors(rng) for _ in range(min(batch, S - s0))]
            vb, tb, nb, kb, pb = self.build_batch(facs)
            verts = vb.astype(np.float32)
            if mesh is None:
                mesh = Mesh(torch.tensor(verts[0], device=dev), torch.tensor(tb, device=dev))
            gs = GeometrySet(mesh, verts)
            ks = [float(k) for k in kb]
            neu = torch.tensor(np.ascontiguousarray(nb), device=dev)
            pts = torch.tensor(pb.astype(np.float32), device=dev)
            lhs, rhs = burton_miller_samples(gs, ks, neu)
            xs, inf = batch_gmres(list(lhs), list(rhs), tol=tol, restart=restart, max_cycles=max_cycles)
            infos.extend(inf)
            p = boundary2potential_samples(gs, ks, neu, torch.stack(xs), pts)
            y[s0:s0 + len(facs), :, 0] = p.abs().float().cpu()  # one copy per batch
            x[s0:s0 + len(facs)] = torch.from_numpy(self.features_batch(facs))
        return x, y, infos


def burton_miller_system(solver, k, neumann):
    """(bm_lhs + M/2, (bm_rhs - beta M/2) psi) with the approximate rules, the
    system BEM_Solver.neumann2dirichlet solves (src/bem/solver.py:118-126)."""
    return solver.burton_miller_system(k, neumann)


def burton_miller_systems(solver, ks, neumanns):
    """burton_miller_system at every wavenumber of ``ks`` on one surface
    (``neumanns``: [len(ks), n]): lhs [K, n, n], rhs [K, n]
    (BEM_Solver.burton_miller_systems: one launch per kernel for all
    wavenumbers)."""
    return solver.burton_miller_systems(ks, neumanns)


def modal_preprocess(obj, mode_num):
    """Surface data of a vibrating solid after ``obj.modal_analysis`` (a
    modal.VibrationObj): per-vertex Neumann data of each mode (the per-triangle
    normal velocity averaged to the vertices) and the wavenumbers, negated as
    the reference stores them (experiments/neuPAT_material_scale/
    preprocess.py:26-60, without the file IO).  Returns the dict
    EditableModalSound takes."""
    from .bem.solver import map_triangle2vertex

    verts = torch.tensor(obj.surf_vertices, dtype=torch.float32, device="cuda")
    tris = torch.tensor(obj.surf_triangles, dtype=torch.int32, device="cuda")
    neumann_vtx = torch.zeros((mode_num, len(verts)), dtype=torch.complex64, device="cuda")
    ks = []
    for i in range(mode_num):
        neumann_tri = torch.tensor(obj.get_triangle_neumann(i), dtype=torch.complex64, device="cuda")
        neumann_vtx[i] = map_triangle2vertex(verts, tris, neumann_tri)
        ks.append(-obj.get_wave_number(i))
    return {"vertices": verts, "triangles": tris, "neumann_vtx": neumann_vtx,
            "ks": torch.tensor(ks, dtype=torch.float32, device="cuda"), "eigenvalues": obj.eigenvalues,
            "modes": obj.modes}


class EditableModalSound:
    """The modal-sound data generator of src/scene.py:301-397: a vibrating
    object's surface modes (``modal_preprocess``), resampled in size and
    material frequency, each mode's Burton-Miller solve and exterior pressure
    at sampled points.  ``data``: the modal_data dict; ``config``: the
    "sample" and "vibration_obj" sections of the reference's config.json.
    ``solve`` solves all modes together (one lockstep batched GMRES over the
    modes' systems); ``solve_loop`` is the reference's one-mode-at-a-time
    loop."""

    def __init__(self, data, config, ffat_res=(64, 32), uniform=False):
        sample_config = config.get("sample", {})
        self.size_base = config.get("vibration_obj", {}).get("size")
        self.vertices_base = data["vertices"]
        self.triangles = data["triangles"]
        self.neumann_vtx = data["neumann_vtx"]
        self.ks_base = data["ks"]
        self.mode_num = len(self.ks_base)
        self.freq_rate = sample_config.get("freq_rate")
        self.size_rate = sample_config.get("size_rate")
        self.bbox_rate = sample_config.get("bbox_rate")
        self.sample_num = sample_config.get("sample_num")
        self.point_num_per_sample = sample_config.get("point_num_per_sample")
        self.ffat_res = ffat_res
        xs = torch.linspace(0, 1, ffat_res[0], device="cuda", dtype=torch.float32)
        ys = torch.linspace(0, 1, ffat_res[1], device="cuda", dtype=torch.float32)
        self.gridx, self.gridy = torch.meshgrid(xs, ys, indexing="ij")
        self.uniform = uniform
        self._solver = None  # the surface's mesh handle, built by the first solve
        if uniform:
            self.point_num_per_sample = ffat_res[0] * ffat_res[1]

    def sample(self, freqK_base=None, sizeK_base=None, generator=None):
        """New size and frequency factors and target points (scene.py:330-383);
        returns the network input [P, 5] = (point factors, size, frequency)."""
        dev = torch.device("cuda")
        self.freqK_base = torch.rand(1, generator=generator).to(dev) if freqK_base is None else freqK_base
        self.sizeK_base = torch.rand(1, generator=generator).to(dev) if sizeK_base is None else sizeK_base
        self.freqK = self.freqK_base * self.freq_rate
        self.sizeK = 1.0 / (1 + self.sizeK_base * (self.size_rate - 1))
        self.vertices = self.vertices_base * self.sizeK
        self.ks = self.ks_base * self.freqK / self.sizeK ** 0.5
        if self.uniform:
            base = torch.zeros(self.ffat_res[0] * self.ffat_res[1], 3, device=dev, dtype=torch.float32)
            base[:, 0] = 0.5
            base[:, 1] = self.gridx.reshape(-1)
            base[:, 2] = self.gridy.reshape(-1)
        else:
            base = torch.rand(self.point_num_per_sample, 3, generator=generator).to(dev)
        self.sample_points_base = base
        rs = (base[:, 0] * (self.bbox_rate - 1) + 1) * self.size_base * 0.7
        theta = base[:, 1] * 2 * np.pi - np.pi
        phi = base[:, 2] * np.pi
        self.trg_points = torch.stack([rs * torch.sin(phi) * torch.cos(theta), rs * torch.sin(phi) * torch.sin(theta),
                                       rs * torch.cos(phi)], dim=-1)
        x = torch.zeros(base.shape[0], 5, dtype=torch.float32)
        x[:, :3] = base.cpu()
        x[:, 3] = self.sizeK_base.cpu()
        x[:, 4] = self.freqK_base.cpu()
        return x

    def solve(self, tol=1e-6, restart=100, max_cycles=10):
        """Pressure of every mode at the target points [modes, P] complex64:
        the modes' Burton-Miller systems (one per wavenumber on the same
        surface) solved together by the lockstep batched GMRES."""
        # one mesh handle across samples: a sample only rescales the surface,
        # so the handle's connectivity (neighbour pairs, tiling) is kept and
        # only its positions are replaced
        if self._solver is None:
            self._solver = BEM_Solver(self.vertices.contiguous(), self.triangles)
        else:
            self._solver.set_vertices(self.vertices.contiguous())
        solver = self._solver
        ks = [float(k) for k in self.ks]
        phi, self.solve_infos = solver.neumann2dirichlet_batch(ks, self.neumann_vtx, tol=tol, restart=restart,
                                                               max_cycles=max_cycles)
        self.potentials = solver.boundary2potential_batch(ks, self.neumann_vtx, phi, self.trg_points).cpu()
        return self.potentials

    def solve_loop(self):
        """The reference's loop (scene.py:384-397): one neumann2dirichlet and
        boundary2potential per mode."""
        solver = BEM_Solver(self.vertices.contiguous(), self.triangles)
        out = []
        for i in range(self.mode_num):
            k = float(self.ks[i])
            d = solver.neumann2dirichlet(k, self.neumann_vtx[i])
            out.append(solver.boundary2potential(k, self.neumann_vtx[i], d, self.trg_points))
        self.potentials = torch.stack(out).cpu()
        return self.potentials


def generate_modal_dataset(modal_sound, sample_num=None, seed=0):
    """x [S, P, 5] (point factors, size, frequency), y [S, P, modes]
    (experiments/neuPAT_material_scale/generate.py:12-37; the reference
    assigns the complex pressure into a float32 tensor, which keeps its real
    part -- kept here)."""
    S = sample_num or modal_sound.sample_num
    gen = torch.Generator().manual_seed(seed)
    P, K = modal_sound.point_num_per_sample, modal_sound.mode_num
    x = torch.zeros(S, P, 5, dtype=torch.float32)
    y = torch.zeros(S, P, K, dtype=torch.float32)
    for q in range(S):
        x[q] = modal_sound.sample(generator=gen)
        y[q] = modal_sound.solve().T.real
    return x, y


def save_dataset(path, x, y):
    """The reference's dataset file: torch.save({"x": x, "y": y}) (generate.py:39)."""
    torch.save({"x": x, "y": y}, path)


__all__ = ["SceneObject", "SceneSampler", "burton_miller_system", "burton_miller_systems", "save_dataset",
           "modal_preprocess", "EditableModalSound", "generate_modal_dataset"]
