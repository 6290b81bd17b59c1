"""CPU oracle of the microstructure generators -- TEST INFRASTRUCTURE ONLY.

A restatement of porerad/geometry.py (the open-cell foam OF, random and
aligned fibers FI / FA, tortuous layers TL; the core, canonical and matched
builders) used as the checker of the device generators; the product never
imports it.  The shape parameters (random draws, periodic neighbour graph,
frames, long-cylinder node lists and disk offsets) are formed with the same
NumPy / SciPy operations as the reference; the N^3 stamping of the short
shapes and the scatter of the long cylinders run in libporerad_oracle.so
(orc_stamp_shapes, orc_stamp_offsets), whose local coordinates follow
OpenBLAS's dgemm rounding (an FMA chain over the three products).

Pinning (tests/test_oracle_geometry.py): the SHA-256 of the phase bytes
listed for every microstructure in data/microstructures.csv, and the
equality of the stamping's local coordinates with NumPy's R.T @ X.
"""

from __future__ import annotations

import ctypes
import hashlib
import math

import numpy as np
from scipy.ndimage import zoom
from scipy.spatial import cKDTree

from . import oracle as _o

CORE_N = 128
PHI_TOL = 3e-3
FIBER_LINES = {"FI": {0.05: 6, 0.10: 10, 0.15: 14}, "FA": {0.05: 6, 0.10: 12, 0.15: 18}}
FIBER_LENGTH = {"FI": 9.0, "FA": 10.0}
FA_THETA_SIGMA = 0.05
FOAM_NODES = {0.05: 24, 0.10: 32, 0.15: 42}
FOAM_NEIGHBOURS = 3
NODE_TO_STRUT = 1.2
TL_LAYERS = {0.05: 9, 0.10: 7, 0.15: 9}
TL_NOISE = 0.12
TL_BASE_N = 192
MATCHED_PAIRS = {1: {"FA": (7, 40), "FI": (0, 12)}, 2: {"FA": (6, 52), "FI": (8, 14)},
                 3: {"FA": (2, 32), "FI": (4, 6)}}
E_X, E_Y, E_Z = np.eye(3)

ELLIPSOID, CYLINDER = 0, 1


def _c():
    L = _o.lib()
    if not hasattr(L, "_geometry_declared"):
        L.orc_stamp_shapes.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p,
                                       ctypes.c_void_p]
        L.orc_stamp_shapes.restype = None
        L.orc_stamp_offsets.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p,
                                        ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_stamp_offsets.restype = None
        L._geometry_declared = True
    return L


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def sha256(phases: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(phases, dtype=np.uint8).tobytes()).hexdigest()


def frame(z, helper):
    """Columns (x, y, z) of a right-handed frame whose third axis is z."""

    if abs(float(np.dot(helper, z))) > 0.9:
        helper = E_Y
    x = np.cross(helper, z)
    x /= np.linalg.norm(x)
    return np.stack([x, np.cross(z, x), z], axis=1)


def long_cylinder_cells(N, r, half_length, R, center):
    """(lattice nodes along the axis, disk offsets) of a cylinder longer than the cell."""

    axis = R @ E_Z
    axis /= float(np.linalg.norm(axis))
    n_steps = max(1, int(math.ceil(2.0 * half_length)))
    t = np.arange(n_steps, dtype=float) - n_steps / 2.0
    nodes = np.unique(np.rint(t[:, None] * axis[None, :] + np.asarray(center, float)[None, :]).astype(np.int64) % N,
                      axis=0)
    Rq = frame(axis / float(np.linalg.norm(axis)), E_X)
    pad = int(math.ceil(r)) + 2
    g = np.arange(-pad, pad + 1, dtype=np.int64)
    offsets = np.stack([a.ravel() for a in np.meshgrid(g, g, g, indexing="ij")], axis=1)
    local = (Rq.T @ offsets.T.astype(float)).T
    disk = ((local[:, 0] / r) ** 2 + (local[:, 1] / r) ** 2 <= 1.0 + 1e-12) & (np.abs(local[:, 2]) <= 0.5 + 1e-12)
    return nodes, offsets[disk]


class Canvas:
    """An N^3 uint8 grid that collects short shapes and long-cylinder stamps."""

    def __init__(self, N):
        self.N = int(N)
        self.phases = np.zeros((self.N,) * 3, np.uint8)
        self.shapes, self.kinds = [], []

    def add(self, center, radii, kind, R):
        N = self.N
        if kind == CYLINDER and float(2.0 * radii[2]) > float(N):
            ci = tuple(int(round(c)) % N for c in center)
            if any(abs(c - float(i)) > 1e-9 for c, i in zip(center, ci)):
                nodes, offs = long_cylinder_cells(N, radii[0], radii[2], R, center)
                shift = (0, 0, 0)
            else:
                nodes, offs = long_cylinder_cells(N, radii[0], radii[2], R, (0.0, 0.0, 0.0))
                shift = ci
            nodes = np.ascontiguousarray(nodes, np.int64)
            offs = np.ascontiguousarray(offs, np.int64)
            sh = np.asarray(shift, np.int64)
            _c().orc_stamp_offsets(_p(self.phases), N, len(nodes), _p(nodes), len(offs), _p(offs), _p(sh))
            return
        self.shapes.append(np.concatenate([np.asarray(center, float), np.asarray(radii, float),
                                           np.asarray(R, float).T.ravel()]))
        self.kinds.append(kind)

    def finish(self):
        if self.shapes:
            params = np.ascontiguousarray(np.stack(self.shapes), np.float64)
            kinds = np.asarray(self.kinds, np.int32)
            _c().orc_stamp_shapes(_p(self.phases), self.N, len(kinds), _p(params), _p(kinds))
            self.shapes, self.kinds = [], []
        return self.phases


def bisect_radius(render, target_phi, r_lo, r_hi, tol=PHI_TOL, max_iter=24):
    lo, hi = float(r_lo), float(r_hi)
    best = None
    for _ in range(int(max_iter)):
        mid = 0.5 * (lo + hi)
        phases = render(mid)
        phi = float(np.count_nonzero(phases == 1) / phases.size)
        err = abs(phi - float(target_phi))
        if best is None or err < best[0]:
            best = (err, mid, phases)
        if err <= tol:
            break
        lo, hi = (mid, hi) if phi < float(target_phi) else (lo, mid)
    return best[1], best[2]


def fiber_axis(rng, aligned):
    if aligned:
        theta = rng.normal(0.0, FA_THETA_SIGMA)
        az = rng.uniform(0.0, 2.0 * math.pi)
        return np.array([math.cos(theta), math.sin(theta) * math.cos(az), math.sin(theta) * math.sin(az)])
    return rng.normal(size=3)


def fiber_lines(N, seed, n_lines, aligned, continuous_centers=False):
    rng = np.random.default_rng(int(seed))
    centers, rotations = [], []
    for _ in range(int(n_lines)):
        c = rng.random(3) * float(N) if continuous_centers else rng.integers(0, N, size=3)
        centers.append(tuple(float(v) for v in c))
        v = fiber_axis(rng, aligned)
        if not aligned:
            v = v / np.linalg.norm(v)
        rotations.append(frame(v / np.linalg.norm(v), E_Z))
    return centers, rotations


def render_fibers(N, centers, rotations, radius, length_factor):
    canvas = Canvas(N)
    half = 0.5 * float(length_factor) * N
    for c, R in zip(centers, rotations):
        canvas.add(c, (radius, radius, half), CYLINDER, R)
    return canvas.finish()


def foam_graph(N, seed, n_nodes, k=FOAM_NEIGHBOURS):
    rng = np.random.default_rng(int(seed))
    points = rng.uniform(0, N, size=(int(n_nodes), 3))
    _, nn = cKDTree(points, boxsize=N).query(points, k=int(k) + 1)
    pairs = set()
    for i, row in enumerate(nn):
        for j in row[1:]:
            if int(i) != int(j):
                pairs.add((min(int(i), int(j)), max(int(i), int(j))))
    return points, sorted(pairs)


def strut(N, p0, p1, r):
    vec = np.asarray(p1, float) - np.asarray(p0, float)
    for a in range(3):
        if vec[a] > N / 2:
            vec[a] -= N
        elif vec[a] < -N / 2:
            vec[a] += N
    L = float(np.linalg.norm(vec))
    center = tuple((np.asarray(p0, float) + 0.5 * vec) % N)
    return center, (r, r, 0.5 * L), frame(vec / L, E_Z)


def render_foam(N, points, edges, strut_radius, node_radius):
    canvas = Canvas(N)
    for p in points:
        canvas.add(tuple(p.tolist()), (node_radius,) * 3, ELLIPSOID, np.eye(3))
    for a, b in edges:
        center, radii, R = strut(N, points[a], points[b], strut_radius)
        canvas.add(center, radii, CYLINDER, R)
    return canvas.finish()


def lamellar_field(N, seed, n_layers, waviness, noise_amp=TL_NOISE):
    rng = np.random.default_rng(int(seed))
    x = np.arange(N, dtype=float)[:, None, None]
    y = np.arange(N, dtype=float)[None, :, None]
    z = np.arange(N, dtype=float)[None, None, :]
    base = np.cos(2.0 * np.pi * n_layers * x / float(N))
    ph1 = rng.uniform(0.0, 2 * np.pi)
    ph2 = rng.uniform(0.0, 2 * np.pi)
    wav = float(waviness) * (0.65 * np.sin(2 * np.pi * y / float(N) + ph1)
                             + 0.35 * np.cos(4 * np.pi * z / float(N) + ph2))
    field = base + 0.35 * np.sin(2 * np.pi * n_layers * (x + wav) / float(N))
    noise = rng.normal(scale=float(noise_amp), size=(N, N, N))
    smooth = noise.copy()
    for axis in range(3):
        for shift in (1, -1):
            smooth = smooth + np.roll(noise, shift, axis)
    return field + smooth / 7.0


def threshold(field, target_phi):
    q = np.quantile(field.ravel(), 1.0 - float(target_phi))
    return (field >= q).astype(np.uint8)


def build(structure, N, target_phi, seed, continuous_centers=False):
    phi = round(float(target_phi), 2)
    if structure in ("FI", "FA"):
        centers, rotations = fiber_lines(N, seed, FIBER_LINES[structure][phi], structure == "FA", continuous_centers)
        lf = FIBER_LENGTH[structure]
        return bisect_radius(lambda r: render_fibers(N, centers, rotations, r, lf), target_phi, 0.1, 4.0)[1]
    if structure == "OF":
        points, edges = foam_graph(N, seed, FOAM_NODES[phi])
        s = max(1.0, float(N) / 48.0)
        return bisect_radius(lambda r: render_foam(N, points, edges, r, float(NODE_TO_STRUT * r)), target_phi,
                             0.2 * s, 4.0 * s)[1]
    if structure == "TL":
        return threshold(lamellar_field(N, seed, TL_LAYERS[phi], 1.4 + 3.0 * float(target_phi)), target_phi)
    raise ValueError(structure)


def canonical(structure, N):
    if structure == "TL":
        field = lamellar_field(TL_BASE_N, 0, TL_LAYERS[0.10], 1.4 + 3.0 * 0.10)
        if N != TL_BASE_N:
            field = zoom(field, float(N) / float(TL_BASE_N), order=3, mode="grid-wrap", grid_mode=True,
                         prefilter=True)
        return threshold(field, 0.10)
    return build(structure, N, 0.10, 0, continuous_centers=True)


def matched(structure, pair, N=CORE_N):
    seed, n_lines = MATCHED_PAIRS[pair][structure]
    aligned = structure == "FA"
    centers = [tuple(float(v) for v in c)
               for c in (np.random.default_rng(seed + 1000003).random((64, 3)) * N)[:n_lines]]
    rng = np.random.default_rng(seed + (3000003 if aligned else 2000003))
    rotations = []
    for _ in range(n_lines):
        v = fiber_axis(rng, aligned)
        rotations.append(frame(v / np.linalg.norm(v), E_Z))
    return bisect_radius(lambda r: render_fibers(N, centers, rotations, r, 9.0), 0.1, 0.12, 7.0, tol=5e-4,
                         max_iter=32)[1]


def generate(row: dict) -> np.ndarray:
    """The microstructure of one row of the microstructure table (id, structure, set, N, target_phi, seed)."""

    s, kind = row["structure"], row["set"]
    if kind == "core":
        return build(s, int(row["N"]), row["target_phi"], int(row["seed"]))
    if kind == "canonical":
        return canonical(s, int(row["N"]))
    if kind == "matched":
        return matched(s, int(row["id"].rsplit("pair", 1)[1]), int(row["N"]))
    raise ValueError(kind)
