"""Synthetic twisted-beam problems and TetGen .node/.ele/handles I/O.

The beam problems are the benchmark workloads: an axis-aligned box of
n_long x n_y x n_z hexahedral cells of edge h along x (the long axis), each cell
split into 6 positive tetrahedra around its main diagonal (Kuhn split, so the
mesh is conforming), every vertex jittered by U(-jitter*h, +jitter*h) per
coordinate from a seeded SplitMix64 stream.  The initial deformation rotates
each cross-section about the beam's centre line by theta_total * t and shifts it
axially by axial * L * t, where t = i / n_long is the cross-section's position
along the beam, so pinned faces keep their exact init positions.

File formats follow the reference loader/saver (samples/utils/MeshIO.hpp:182-362):
header "<n> 3 0 0" / "<n> 4 0", one "<idx> ..." line per entry, 0- or 1-based
indices (detected from the first line), and a handles file of whitespace
separated vertex indices (samples/tests/quasi.cpp:146-159).
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field

import numpy as np

_MASK = np.uint64(0xFFFFFFFFFFFFFFFF)


class SplitMix64:
    """SplitMix64 (Steele, Lea & Flood 2014); identical streams on every host."""

    def __init__(self, seed: int):
        self.state = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)

    def next_u64(self, n: int) -> np.ndarray:
        with np.errstate(over="ignore"):
            inc = np.uint64(0x9E3779B97F4A7C15)
            steps = np.arange(1, n + 1, dtype=np.uint64)
            z = self.state + steps * inc
            self.state = self.state + np.uint64(n) * inc
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
        return z

    def uniform(self, n: int) -> np.ndarray:
        """n doubles in [0, 1) from the top 53 bits."""
        return (self.next_u64(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


# The 6 tetrahedra of a unit cube with corners indexed by bits (dx | dy<<1 | dz<<2),
# one per axis permutation, all containing the 0 -> 7 diagonal.  Orientation is
# fixed below so every tet has positive volume.
def _kuhn_tets() -> np.ndarray:
    import itertools
    tets = []
    for perm in itertools.permutations(range(3)):
        c = 0
        verts = [c]
        for ax in perm:
            c |= 1 << ax
            verts.append(c)
        p = [np.array([(v >> 0) & 1, (v >> 1) & 1, (v >> 2) & 1], float) for v in verts]
        det = np.linalg.det(np.stack([p[1] - p[0], p[2] - p[0], p[3] - p[0]], axis=1))
        if det < 0:
            verts[1], verts[2] = verts[2], verts[1]
        tets.append(verts)
    return np.array(tets, dtype=np.int64)


@dataclass
class Problem:
    """A complete solver input: rest/init positions, tets, pins and material."""

    name: str
    rest: np.ndarray            # (n_v, 3) float64
    init: np.ndarray            # (n_v, 3) float64
    tets: np.ndarray            # (n_e, 4) int32
    pins: np.ndarray            # sorted int32 vertex indices
    model: str = "nh"           # "nh" or "arap"
    youngs: float = 1e7
    poisson: float = 0.45
    youngs_per_elem: np.ndarray | None = None   # (n_e,) for heterogeneous material
    settings: dict = field(default_factory=dict)

    @property
    def n_verts(self) -> int:
        return self.rest.shape[0]

    @property
    def n_elems(self) -> int:
        return self.tets.shape[0]

    def lame(self):
        e, nu = self.youngs, self.poisson
        return e / (2.0 * (1.0 + nu)), e * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))

    def lame_per_elem(self):
        if self.youngs_per_elem is None:
            return None
        e, nu = self.youngs_per_elem, self.poisson
        return e / (2.0 * (1.0 + nu)), e * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))


def beam(n_long: int, n_y: int, n_z: int, *, h: float = 0.1, jitter: float = 0.01,
         seed: int = 1, twist_deg: float = 90.0, axial: float = 0.0,
         pin_far_end: bool = True, voronoi_regions: int = 0,
         youngs_choices=(1e5, 1e7)):
    """Generate a twisted beam.  Returns (rest, init, tets, pins, youngs_per_elem)."""
    rng = SplitMix64(seed)
    nx, ny, nz = n_long + 1, n_y + 1, n_z + 1
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    vid = (i * ny + j) * nz + k
    rest = np.stack([i * h, j * h, k * h], axis=-1).reshape(-1, 3).astype(np.float64)
    n_v = rest.shape[0]
    rest += (rng.uniform(3 * n_v).reshape(n_v, 3) * 2.0 - 1.0) * (jitter * h)

    # tets, cell-major so neighbouring elements share vertices in memory
    ci, cj, ck = np.meshgrid(np.arange(n_long), np.arange(n_y), np.arange(n_z), indexing="ij")
    ci, cj, ck = ci.ravel(), cj.ravel(), ck.ravel()
    corner = np.empty((ci.size, 8), dtype=np.int64)
    for b in range(8):
        dx, dy, dz = b & 1, (b >> 1) & 1, (b >> 2) & 1
        corner[:, b] = vid[ci + dx, cj + dy, ck + dz]
    kt = _kuhn_tets()
    tets = corner[:, kt].reshape(-1, 4).astype(np.int32)

    t = (np.arange(n_v) // (ny * nz)).astype(np.float64) / n_long   # cross-section index / n_long
    length = n_long * h
    yc, zc = 0.5 * n_y * h, 0.5 * n_z * h
    theta = np.deg2rad(twist_deg) * t
    c, s = np.cos(theta), np.sin(theta)
    y0, z0 = rest[:, 1] - yc, rest[:, 2] - zc
    init = rest.copy()
    init[:, 1] = yc + c * y0 - s * z0
    init[:, 2] = zc + s * y0 + c * z0
    init[:, 0] = rest[:, 0] + axial * length * t

    sec = np.arange(n_v) // (ny * nz)
    pin_mask = sec == 0
    if pin_far_end:
        pin_mask |= sec == n_long
    pins = np.nonzero(pin_mask)[0].astype(np.int32)

    ype = None
    if voronoi_regions > 0:
        seeds = rng.uniform(3 * voronoi_regions).reshape(voronoi_regions, 3) * np.array(
            [length, n_y * h, n_z * h])
        pick = rng.next_u64(voronoi_regions) & np.uint64(1)
        region_e = np.where(pick == 1, youngs_choices[1], youngs_choices[0]).astype(np.float64)
        centres = (np.stack([ci, cj, ck], axis=1).astype(np.float64) + 0.5) * h
        d2 = ((centres[:, None, :] - seeds[None, :, :]) ** 2).sum(-1)
        cell_region = np.argmin(d2, axis=1)
        ype = np.repeat(region_e[cell_region], 6)
    return rest, init, tets, pins, ype


def config_problem(cfg: int) -> Problem:
    """The five benchmark configurations (seed = configuration number)."""
    if cfg == 1:
        r, i0, t, p, _ = beam(20, 4, 4, seed=1, twist_deg=90.0, pin_far_end=False)
        return Problem("C1 beam 4x4x20 ARAP 90deg", r, i0, t, p, model="arap",
                       settings=dict(reweighting=0, max_admm_iters=100))
    if cfg in (2, 4):
        r, i0, t, p, _ = beam(160, 16, 16, seed=cfg, twist_deg=360.0)
        return Problem(f"C{cfg} beam 16x16x160 NH 360deg", r, i0, t, p, model="nh", youngs=1e6,
                       settings=dict(reweighting=1, lbfgs_m=5, max_admm_iters=500))
    if cfg == 3:
        r, i0, t, p, _ = beam(240, 32, 32, seed=3, twist_deg=180.0, axial=-0.4)
        return Problem("C3 beam 32x32x240 NH push+180deg", r, i0, t, p, model="nh", youngs=1e6,
                       settings=dict(reweighting=1, lbfgs_m=5, max_admm_iters=500))
    if cfg == 5:
        r, i0, t, p, ype = beam(320, 40, 40, seed=5, twist_deg=360.0, axial=-0.2,
                                voronoi_regions=64)
        return Problem("C5 beam 40x40x320 NH voronoi 360deg+20%", r, i0, t, p, model="nh",
                       youngs=1e6, youngs_per_elem=ype,
                       settings=dict(reweighting=1, lbfgs_m=5, max_admm_iters=500))
    raise ValueError(f"unknown configuration {cfg}")


def small_voronoi_beam(n_long=24, n_y=4, n_z=4, *, regions=8, seed=21, twist_deg=240.0, axial=-0.15,
                       **settings) -> Problem:
    """A small heterogeneous beam: per-region E in {1e5, 1e7} over `regions` seeded Voronoi
    cells, twisted and pushed like the 3.07M-tet benchmark beam."""
    r, i0, t, p, ype = beam(n_long, n_y, n_z, seed=seed, twist_deg=twist_deg, axial=axial, voronoi_regions=regions)
    return Problem(f"voronoi beam {n_y}x{n_z}x{n_long}", r, i0, t, p, model="nh", youngs=1e6, youngs_per_elem=ype,
                   settings=dict(settings))


def small_beam(n_long=8, n_y=2, n_z=2, *, twist_deg=90.0, axial=0.0, seed=7, model="nh",
               youngs=1e7, pin_far_end=True, **settings) -> Problem:
    r, i0, t, p, _ = beam(n_long, n_y, n_z, seed=seed, twist_deg=twist_deg, axial=axial,
                          pin_far_end=pin_far_end)
    return Problem(f"beam {n_y}x{n_z}x{n_long}", r, i0, t, p, model=model, youngs=youngs,
                   settings=dict(settings))


# --------------------------------------------------------------------------------------
# TetGen-style files
def save_elenode(path_noext: str, verts: np.ndarray, tets: np.ndarray) -> None:
    """Write <path>.node and <path>.ele; coordinates with 17 significant digits."""
    with open(path_noext + ".ele", "w") as f:
        f.write(f"{tets.shape[0]} 4 0\n")
        for n, t in enumerate(tets):
            f.write(f"\t{n} {t[0]} {t[1]} {t[2]} {t[3]}\n")
    with open(path_noext + ".node", "w") as f:
        f.write(f"{verts.shape[0]} 3 0 0\n")
        for n, v in enumerate(verts):
            f.write(f"\t{n} {v[0]:.17g} {v[1]:.17g} {v[2]:.17g}\n")


def load_elenode(path_noext: str):
    """Read <path>.node/.ele with the reference's 1-based detection."""
    def _rows(fn, ncol):
        with open(fn) as f:
            n = int(f.readline().split()[0])
            rows = []
            for _ in range(n):
                rows.append(f.readline().split()[:ncol])
        return rows
    base = os.path.splitext(path_noext)[0] if path_noext.endswith((".node", ".ele")) else path_noext
    er = _rows(base + ".ele", 5)
    idx = np.array([int(r[0]) for r in er])
    tv = np.array([[int(x) for x in r[1:5]] for r in er], dtype=np.int64)
    if idx[0] == 1:
        idx, tv = idx - 1, tv - 1
    tets = np.empty_like(tv)
    tets[idx] = tv
    nr = _rows(base + ".node", 4)
    nidx = np.array([int(r[0]) for r in nr])
    xyz = np.array([[float(x) for x in r[1:4]] for r in nr])
    if nidx[0] == 1:
        nidx = nidx - 1
    verts = np.empty_like(xyz)
    verts[nidx] = xyz
    return verts, tets.astype(np.int32)


def save_handles(path: str, pins: np.ndarray) -> None:
    with open(path, "w") as f:
        for p in pins:
            f.write(f"{int(p)}\n")


def load_handles(path: str) -> np.ndarray:
    with open(path) as f:
        return np.array([int(x) for x in f.read().split()], dtype=np.int32)


def write_problem(prob: Problem, directory: str) -> dict:
    """Write rest/init/handles files; returns the paths."""
    os.makedirs(directory, exist_ok=True)
    rest = os.path.join(directory, "rest")
    init = os.path.join(directory, "init")
    hnd = os.path.join(directory, "handles.txt")
    save_elenode(rest, prob.rest, prob.tets)
    save_elenode(init, prob.init, prob.tets)
    save_handles(hnd, prob.pins)
    return dict(rest=rest, init=init, handles=hnd)
