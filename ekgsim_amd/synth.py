"""Seeded synthetic hearts, leads and parameter vectors.

The reference ships one real heart (testRun/model_24.matrix) and a 2D ring
generator (simlib/simulator.h:600-644).  The benchmark workloads use a 3D
generator instead: myocardium is the shell between two nested axis-aligned
ellipsoids, split into endo/mid/epi thirds (layers 1/2/3) by a transmural
coordinate, with a per-voxel conduction delay of 1-3 steps and excitation
seeds on the endocardial surface.  Coordinates follow the reference's storage
order (i0, i1, i2) = (z, y, x) with i2 fastest (Hypermatrix.h:652-661).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

# Base AP coefficients of the shipped settings (testRun/simulator.ini:48-50)
# and the k min / k max bounds (simulator.ini:52-53).
BASE_AP = (0.0, 2.5, 100.0, 0.90, 0.1, 0.0010, 0.10, 0.10, 250.0)
K_MIN = (0.0, 1.5, 100.0, 0.85, 0.05, 0.0003, 0.01, 0.01, 300.0)
K_MAX = (0.0, 3.5, 100.0, 0.95, 0.2, 0.0010, 0.10, 0.10, 450.0)
FREE_K = (5, 6, 7, 8)
# Sampling bounds of the free coefficients (k5, k6, k7, k8) and of the
# electrode displacements, as quoted with the README's 16-gene vector
# (README.md:94-104).
FREE_LO = (0.0003, 0.01, 0.01, 200.0)
FREE_HI = (0.0010, 0.10, 0.10, 400.0)
DISP_LO, DISP_HI = -50.0, 50.0

# (outer semi-axes, inner semi-axes) in voxels along (x, y, z).
ELLIPSOIDS = {
    48: ((20, 16, 22), (14, 10, 16)),
    128: ((54, 42, 58), (38, 27, 42)),
    256: ((108, 84, 116), (76, 54, 84)),
}


@dataclasses.dataclass
class Heart:
    """A voxel heart: layer (uint8, 0 = empty), seed flags, per-voxel delay."""

    layer: np.ndarray          # (n0, n1, n2) uint8
    seed: np.ndarray           # (n0, n1, n2) uint8, 1 = excitation start
    delay: np.ndarray          # (n0, n1, n2) uint8, steps, 0 outside myocardium
    n_layers: int = 3

    @property
    def dims(self):
        return tuple(int(d) for d in self.layer.shape)

    @property
    def n_myo(self) -> int:
        return int(np.count_nonzero(self.layer))


def ellipsoid_axes(n: int):
    """Semi-axes for an n^3 grid: the tabulated ones, else the 128^3 ones
    scaled by n/128."""
    if n in ELLIPSOIDS:
        return ELLIPSOIDS[n]
    (o, i) = ELLIPSOIDS[128]
    s = n / 128.0
    return tuple(v * s for v in o), tuple(v * s for v in i)


def make_heart(n: int, seed: int = 0, n_seeds: int = 5, dims=None, n_layers: int = 3) -> Heart:
    """Ellipsoid-shell heart on an n^3 grid (or dims = (n0, n1, n2)) whose
    myocardium is split into n_layers transmural layers of equal thickness."""
    if dims is None:
        dims = (n, n, n)
    n0, n1, n2 = dims
    (ax, ay, az), (bx, by, bz) = ellipsoid_axes(n)
    z = np.arange(n0) - (n0 - 1) / 2.0
    y = np.arange(n1) - (n1 - 1) / 2.0
    x = np.arange(n2) - (n2 - 1) / 2.0
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    q_out = (X / ax) ** 2 + (Y / ay) ** 2 + (Z / az) ** 2
    q_in = (X / bx) ** 2 + (Y / by) ** 2 + (Z / bz) ** 2
    myo = (q_out < 1.0) & (q_in > 1.0)
    r_out = np.sqrt(q_out)
    r_in = np.sqrt(q_in)
    with np.errstate(divide="ignore", invalid="ignore"):
        s = (r_in - 1.0) / ((r_in - 1.0) + (1.0 - r_out))
    lay = np.zeros(dims, np.uint8)
    lay[myo] = (1 + np.minimum(n_layers - 1, np.floor(n_layers * s[myo]))).astype(np.uint8)
    rng = np.random.default_rng(seed)
    delay = np.zeros(dims, np.uint8)
    delay[myo] = rng.integers(1, 4, size=int(myo.sum()), dtype=np.uint8)
    # endocardial voxels: layer 1 with an empty voxel among its 26 neighbours
    pad = np.pad(lay, 1)
    empty_nb = np.zeros(dims, bool)
    for d0 in (-1, 0, 1):
        for d1 in (-1, 0, 1):
            for d2 in (-1, 0, 1):
                if d0 == d1 == d2 == 0:
                    continue
                sl = pad[1 + d0:1 + d0 + n0, 1 + d1:1 + d1 + n1, 1 + d2:1 + d2 + n2]
                empty_nb |= sl == 0
    endo = np.flatnonzero(((lay == 1) & empty_nb).ravel())
    seeds = np.zeros(dims, np.uint8)
    if len(endo):
        pick = rng.choice(endo, size=min(n_seeds, len(endo)), replace=False)
        seeds.ravel()[pick] = 1
    return Heart(layer=lay, seed=seeds, delay=delay, n_layers=n_layers)


def lead_positions(n: int, n_leads: int, dims=None) -> np.ndarray:
    """n_leads points of a Fibonacci lattice on the sphere of radius
    1.5 * n/2 about the grid centre, in voxel units relative to the model
    corner, stored (z, y, x) like the reference's measuring points."""
    if dims is None:
        dims = (n, n, n)
    c = np.array([(d - 1) / 2.0 for d in dims])  # (z, y, x)
    r = 1.5 * n / 2.0
    pts = np.zeros((n_leads, 3))
    golden = math.pi * (3.0 - math.sqrt(5.0))
    for i in range(n_leads):
        yy = 1.0 - 2.0 * (i + 0.5) / n_leads
        rad = math.sqrt(max(0.0, 1.0 - yy * yy))
        th = golden * i
        px, py, pz = math.cos(th) * rad, yy, math.sin(th) * rad
        pts[i] = c + r * np.array([pz, py, px])
    return pts


def random_params(n_layers: int, n_leads: int, seed: int):
    """A parameter vector in the reference's -sim layout: 4 free k per
    layer (endo, mid, epi) followed by (du, dv) per lead."""
    rng = np.random.default_rng(seed)
    ks = []
    for _ in range(n_layers):
        ks.extend(rng.uniform(FREE_LO, FREE_HI))
    disp = rng.uniform(DISP_LO, DISP_HI, size=2 * n_leads)
    return np.concatenate([np.asarray(ks), disp])


def layer_k_from_params(x, n_layers: int = 3, base=BASE_AP):
    """Layer coefficient table (n_layers, 9) from the free genes, for the
    3-layer endo-mid-epi case where every layer starts from the same base AP
    (sim.cpp:843-875 with identical base APs)."""
    k = np.tile(np.asarray(base, np.float64), (n_layers, 1))
    for l in range(n_layers):
        k[l, list(FREE_K)] = x[4 * l:4 * l + 4]
    return k


def displacements(x, n_layers: int, n_leads: int) -> np.ndarray:
    return np.asarray(x[4 * n_layers:4 * n_layers + 2 * n_leads], np.float64).reshape(n_leads, 2)


def lead_basis(p):
    """U/V displacement basis of one lead (sim_lib.h:181-193) in FP64."""
    p = np.asarray(p, np.float64)
    u = np.array([0.0 * p[2] - 1.0 * p[1], 1.0 * p[0] - 0.0 * p[2], 0.0 * p[1] - 0.0 * p[0]])
    v = np.array([u[1] * p[2] - u[2] * p[1], u[2] * p[0] - u[0] * p[2], u[0] * p[1] - u[1] * p[0]])
    u = -u

    def nrm(w):
        l = math.sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2])
        return w * (1.0 / l) if l != 0 else w * 0.0

    return nrm(u), nrm(v)


def move_leads(pos, disp):
    """Displaced lead positions (sim_lib.h:197-208)."""
    out = np.empty_like(np.asarray(pos, np.float64))
    for i, p in enumerate(pos):
        U, V = lead_basis(p)
        out[i] = (p + U * disp[i][0]) + V * disp[i][1]
    return out


def write_matrix(path, layer, seed=None):
    """Write a heart in the reference's .matrix text format
    (matrix.h:133-248): comment line, '3d X x Y x Z space', then rows of i2
    values, one block per i0; a negative value marks a seed."""
    n0, n1, n2 = layer.shape
    vals = layer.astype(np.int64)
    if seed is not None:
        vals = np.where(seed > 0, -vals, vals)
    with open(path, "w") as f:
        f.write("# synthetic heart\n")
        f.write(f"3d {n2} x {n1} x {n0} space\n")
        for i0 in range(n0):
            for i1 in range(n1):
                f.write(" " + " ".join(str(int(v)) for v in vals[i0, i1]) + "\n")
            f.write("\n")


def write_transfer(path, t):
    t = np.asarray(t, np.float64)
    rows, cols = t.shape
    with open(path, "w") as f:
        f.write("Conduction matrix\n")
        f.write(f"2d {cols} x {rows} tab\n")
        for r in range(rows):
            f.write("\t".join(repr(float(v)) for v in t[r]) + "\n")


def write_points(path, pos):
    """Measuring points file: one 'x, y, z' line per lead (simulator.cpp:369-390)."""
    with open(path, "w") as f:
        for p in pos:
            f.write(f"{float(p[2])!r}, {float(p[1])!r}, {float(p[0])!r}\n")


def layered_transfer(n_layers: int = 3, size: int = 12, diag: float = 1.0, off: float = 1.5):
    """A conduction matrix of at least 12 x 12 (sim.cpp:371-374 checks the
    size against the default of 12 layers before the shape is loaded): row
    and column 0 are -1 (no tissue) like testRun/conduction_24.matrix."""
    size = max(size, n_layers + 1)
    t = np.full((size, size), -1.0)
    for a in range(1, size):
        for b in range(1, size):
            t[a, b] = diag if a == b else off * abs(a - b)
    return t
