"""Deterministic synthetic microstructures for the benchmark workloads.

These generators are this package's own (the reference builds its geometries
with porerad/geometry.py, which is not part of the radiation path); they only
produce inputs in the reference's format: uint8 arrays (nx, ny, nz) with
1 = solid, 0 = pore, periodic in every direction.

  random_fibers   straight cylinders of given radius, isotropic uniform axes
                  and uniform centres, added until the solid fraction is met
  voronoi_foam    open-cell foam: struts along the edges of the periodic
                  Voronoi tessellation of uniform seeds, strut radius bisected
                  to the solid fraction
  aligned_fibers  cylinders along z with Poisson-disk centres
  layered_plates  plates normal to x with square holes (tortuous layers)
"""

from __future__ import annotations

import numpy as np


def _centres(n: int) -> np.ndarray:
    return np.arange(n, dtype=np.float64) + 0.5


def random_fibers(N: int, solid_fraction: float, radius: float, seed: int, max_fibers: int = 100000) -> np.ndarray:
    """Straight cylinders (length N, minimum-image periodic) until the solid fraction is reached.

    A voxel centre x belongs to a fibre (centre c, unit axis a) when the
    minimum-image offset v = x - c lies within ``radius`` of the segment
    t a, |t| <= N/2.  Only the voxels of the fibre's tube are visited: in each
    layer across the axis's dominant component d the tube's section is an
    ellipse of semi-axes radius and radius / |a_d| around the axis, so a window
    of that half-width (plus one voxel) in the two other coordinates holds
    every voxel the fibre can reach.
    """

    rng = np.random.default_rng(seed)
    c = _centres(N)
    solid = np.zeros((N, N, N), dtype=bool)
    half = N / 2.0
    r2 = radius * radius
    layers = np.arange(N)
    target = solid_fraction * solid.size
    filled = 0
    for _ in range(max_fibers):
        if filled >= target:
            break
        centre = rng.uniform(0.0, N, size=3)
        axis = rng.normal(size=3)
        axis /= np.linalg.norm(axis)
        d = int(np.argmax(np.abs(axis)))
        e1, e2 = (d + 1) % 3, (d + 2) % 3
        w = int(np.ceil(radius / abs(axis[d]))) + 1
        offs = np.arange(-w, w + 1)
        vd = c - centre[d]
        vd -= N * np.round(vd / N)
        idx = [None, None, None]
        idx[d] = np.broadcast_to(layers[:, None, None], (N, offs.size, offs.size))
        for e, shape in ((e1, (N, offs.size, 1)), (e2, (N, 1, offs.size))):
            mid = centre[e] + vd * (axis[e] / axis[d]) - 0.5  # axis position as a fractional index
            i = (np.floor(mid)[:, None] + offs[None, :]).astype(np.int64) % N
            idx[e] = np.broadcast_to(i.reshape(shape), (N, offs.size, offs.size))
        v = [c[idx[k]] - centre[k] for k in range(3)]
        for k in range(3):
            v[k] -= N * np.round(v[k] / N)
        t = np.clip(v[0] * axis[0] + v[1] * axis[1] + v[2] * axis[2], -half, half)
        dist2 = (v[0] - t * axis[0]) ** 2 + (v[1] - t * axis[1]) ** 2 + (v[2] - t * axis[2]) ** 2
        hit = dist2 <= r2
        flat = np.unique(np.ravel_multi_index((idx[0][hit], idx[1][hit], idx[2][hit]), solid.shape))
        flat = flat[~solid.ravel()[flat]]
        solid.ravel()[flat] = True
        filled += flat.size
    return solid.astype(np.uint8)


def _segment_distance_field(N: int, segments: np.ndarray, reach: float) -> np.ndarray:
    """Distance from every voxel centre to the nearest segment, capped at ``reach``."""

    c = _centres(N)
    best = np.full((N, N, N), np.inf)
    for a, b in segments:
        lo = np.floor(np.minimum(a, b) - reach).astype(int)
        hi = np.ceil(np.maximum(a, b) + reach).astype(int)
        lo = np.maximum(lo, 0)
        hi = np.minimum(hi, N)
        if np.any(hi <= lo):
            continue
        xs, ys, zs = c[lo[0]:hi[0]], c[lo[1]:hi[1]], c[lo[2]:hi[2]]
        X, Y, Z = np.meshgrid(xs, ys, zs, indexing="ij")
        ab = b - a
        L2 = float(ab @ ab)
        px, py, pz = X - a[0], Y - a[1], Z - a[2]
        t = (px * ab[0] + py * ab[1] + pz * ab[2]) / L2 if L2 > 0 else np.zeros_like(px)
        t = np.clip(t, 0.0, 1.0)
        d = np.sqrt((px - t * ab[0]) ** 2 + (py - t * ab[1]) ** 2 + (pz - t * ab[2]) ** 2)
        view = best[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]]
        np.minimum(view, d, out=view)
    return best


def voronoi_foam(N: int, n_seeds: int, solid_fraction: float, seed: int, reach: float = 10.0) -> np.ndarray:
    """Open-cell foam from the edges of the periodic Voronoi tessellation of ``n_seeds`` uniform points."""

    from scipy.spatial import Voronoi

    rng = np.random.default_rng(seed)
    pts = rng.uniform(0.0, N, size=(n_seeds, 3))
    shifts = np.array([(i, j, k) for i in (-1, 0, 1) for j in (-1, 0, 1) for k in (-1, 0, 1)], dtype=np.float64)
    images = (pts[None, :, :] + N * shifts[:, None, :]).reshape(-1, 3)
    vor = Voronoi(images)
    verts = vor.vertices
    edges = set()
    for ridge in vor.ridge_vertices:
        if -1 in ridge:
            continue
        for u, v in zip(ridge, ridge[1:] + ridge[:1]):
            edges.add((min(u, v), max(u, v)))
    segs = []
    lo, hi = -reach, N + reach
    for u, v in sorted(edges):
        a, b = verts[u], verts[v]
        if np.all(np.minimum(a, b) < hi) and np.all(np.maximum(a, b) > lo):
            segs.append((a, b))
    dist = _segment_distance_field(N, segs, reach)
    target = int(round(solid_fraction * N**3))
    r_lo, r_hi = 0.0, reach
    for _ in range(60):
        r = 0.5 * (r_lo + r_hi)
        if np.count_nonzero(dist <= r) < target:
            r_lo = r
        else:
            r_hi = r
    return (dist <= r_hi).astype(np.uint8)


def aligned_fibers(N: int, solid_fraction: float, radius: float, seed: int, gap: float = 1.0) -> np.ndarray:
    """Cylinders along z with Poisson-disk centres (minimum spacing 2 r + gap), periodic in x and y."""

    rng = np.random.default_rng(seed)
    c = _centres(N)
    X, Y = np.meshgrid(c, c, indexing="ij")
    section = np.zeros((N, N), dtype=bool)
    centres: list[np.ndarray] = []
    min_d = 2.0 * radius + gap
    attempts = 0
    while section.mean() < solid_fraction and attempts < 200000:
        attempts += 1
        p = rng.uniform(0.0, N, size=2)
        ok = True
        for q in centres:
            d = p - q
            d -= N * np.round(d / N)
            if d @ d < min_d * min_d:
                ok = False
                break
        if not ok:
            continue
        centres.append(p)
        dx = X - p[0]
        dy = Y - p[1]
        dx -= N * np.round(dx / N)
        dy -= N * np.round(dy / N)
        section |= dx * dx + dy * dy <= radius * radius
    return np.broadcast_to(section[:, :, None], (N, N, N)).astype(np.uint8).copy()


def layered_plates(N: int, solid_fraction: float, seed: int, thickness: int = 2, hole: int = 6,
                   hole_fraction: float = 0.15) -> np.ndarray:
    """Plates normal to x, ``thickness`` cells thick, with random square holes covering ``hole_fraction``."""

    rng = np.random.default_rng(seed)
    period = max(thickness + 1, int(round(thickness * (1.0 - hole_fraction) / solid_fraction)))
    out = np.zeros((N, N, N), dtype=np.uint8)
    for start in range(period // 2, N - thickness + 1, period):
        plate = np.ones((N, N), dtype=bool)
        while 1.0 - plate.mean() < hole_fraction:
            j, k = rng.integers(0, N, size=2)
            jj = (j + np.arange(hole)) % N
            kk = (k + np.arange(hole)) % N
            plate[np.ix_(jj, kk)] = False
        out[start:start + thickness] = plate[None, :, :]
    return out


# ---------------------------------------------------------------- named workloads
# The geometries of the benchmark and of the full-size tests, x-oriented as
# generated (view_factors.orient_along_x turns them to another axis).

def fibre_n64() -> np.ndarray:
    """Random fibres N=64, solid fraction 0.10, radius 3, seed 0 (traced with 200 rays/patch, seed 1)."""

    return random_fibers(64, 0.10, 3.0, 0)


def foam_n128() -> np.ndarray:
    """Voronoi-edge foam N=128 from 64 seeds, solid fraction 0.10, seed 2 (1000 rays/patch, seed 3)."""

    return voronoi_foam(128, 64, 0.10, 2)


def aligned_fibre_n128() -> np.ndarray:
    """Cylinders of radius 4 along z with Poisson-disk centres, N=128, solid fraction 0.10, seed 0."""

    return aligned_fibers(128, 0.10, 4.0, 0)


MIXED_SIZES = (96, 128, 160, 192)


def mixed_geometries(sizes=MIXED_SIZES) -> list[tuple[str, np.ndarray]]:
    """Foam, random fibres, aligned fibres and perforated plates (2 cells thick, 6-cell holes over 15 %
    of each plate) at each N of ``sizes``: 16 uneven geometries by default, traced with seeds 10, 11, ...
    in this order.  Foam cells keep their size: 64 Voronoi seeds per 128^3."""

    families = [
        ("foam", lambda N: voronoi_foam(N, max(8, int(round(64 * (N / 128.0) ** 3))), 0.10, 2)),
        ("random_fibre", lambda N: random_fibers(N, 0.10, 3.0, 0)),
        ("aligned_fibre", lambda N: aligned_fibers(N, 0.10, 4.0, 0)),
        ("perforated_plates", lambda N: layered_plates(N, 0.10, 0, thickness=2, hole=6, hole_fraction=0.15)),
    ]
    return [(f"{name}{N}", fn(N)) for N in sizes for name, fn in families]
