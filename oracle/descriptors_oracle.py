"""CPU oracle of the morphology descriptors -- TEST INFRASTRUCTURE ONLY.

A NumPy restatement of porerad/descriptors.py (scalar descriptors 38-79,
directional descriptors 82-180) used as the checker of the device
descriptors; the product never imports it.  Cluster labelling calls
scipy.ndimage.label with a full 3x3x3 structure, the routine the reference
calls (scipy 1.15.3, requirements.txt:2).

Pinning (tests/test_oracle_descriptors.py):
  scalar counts ....... realized_phi, interfacial_area_density_per_cell,
                        mean_beam_length_cells and mean_pore_chord_x/y/z of
                        data/microstructures.csv for every fixture geometry
                        (exact); pure-Python line walks on small grids
  connectivity ........ breadth-first search over the 26 neighbours on small
                        grids; every accepted OF core geometry percolates
                        along x, y, z with a largest cluster >= 0.99 of the
                        pore (the acceptance rule, porerad/geometry.py:256-275)
  periodic chords ..... pure-Python periodic line walks on small grids (order
                        and joined end runs); the chord lengths of each axis
                        add up to the pore cells
  lineal path ......... L(1) = porosity; a single chord of length c gives
                        (c - l + 1) / (lines N); an all-pore line gives 1
  summary rows ........ parallel plates of known spacing (closed-form chord
                        mean, quantiles and face counts)
"""

from __future__ import annotations

import numpy as np
from scipy import ndimage

QUANTILES = (0.5, 0.75, 0.9, 0.95, 0.99)
LINEAL_SAMPLES = (1, 2, 4, 8, 16, 32, 64, 128)


def axis_lines(mask: np.ndarray, axis: int) -> np.ndarray:
    """(lines, length) view: lines in row-major order of the two other axes."""

    m = np.asarray(mask, bool)
    order = [a for a in range(3) if a != axis] + [axis]
    t = np.transpose(m, order)
    return t.reshape(-1, m.shape[axis])


def run_table(rows: np.ndarray):
    """(row, start, end) of every maximal True run, in row-major order; end is exclusive."""

    M, L = rows.shape
    edge = np.zeros((M, L + 1), np.int8)
    edge[:, :L] += rows
    edge[:, 1:] -= rows
    # +1 where a run starts (this cell True, previous False), -1 one past its end
    r_start, c_start = np.nonzero(edge == 1)
    r_end, c_end = np.nonzero(edge == -1)
    assert np.array_equal(r_start, r_end)
    return r_start, c_start, c_end


def box_faces(solid: np.ndarray) -> np.ndarray:
    s = np.asarray(solid, bool)
    return np.array([int(np.sum(np.diff(s, axis=a) != 0)) for a in range(3)], np.int64)


def periodic_faces(solid: np.ndarray) -> np.ndarray:
    s = np.asarray(solid, bool)
    out = []
    for a in range(3):
        nxt = np.concatenate([np.take(s, np.arange(1, s.shape[a]), axis=a), np.take(s, [0], axis=a)], axis=a)
        out.append(int(np.sum(s != nxt)))
    return np.array(out, np.int64)


def pore_runs(pore: np.ndarray) -> np.ndarray:
    return np.array([run_table(axis_lines(pore, a))[0].size for a in range(3)], np.int64)


def clusters(pore: np.ndarray):
    """(percolation per axis, largest cluster size, cluster count) of the 26-connected pore."""

    labels, count = ndimage.label(np.asarray(pore, bool), structure=np.ones((3, 3, 3), bool))
    sizes = np.bincount(labels.ravel(), minlength=count + 1)
    sizes[0] = 0
    perc = []
    for a in range(3):
        lo = set(np.unique(np.take(labels, 0, axis=a)).tolist()) - {0}
        hi = set(np.unique(np.take(labels, labels.shape[a] - 1, axis=a)).tolist()) - {0}
        perc.append(bool(lo & hi))
    return tuple(perc), int(sizes.max()), int(count)


def scalar_counts(phases: np.ndarray) -> dict:
    p = np.asarray(phases)
    solid, pore = p == 1, p == 0
    perc, largest, count = clusters(pore)
    return {"cells": int(p.size), "solid": int(solid.sum()), "pore": int(pore.sum()),
            "box_faces": box_faces(solid), "pore_runs": pore_runs(pore), "periodic_faces": periodic_faces(p == 0),
            "percolates": perc, "largest_cluster": largest, "clusters": count}


def scalar_descriptors(phases: np.ndarray) -> dict:
    c = scalar_counts(phases)
    area = int(c["box_faces"].sum())
    chords = tuple(float(np.float64(c["pore"]) / np.float64(r)) if r else float("nan") for r in c["pore_runs"])
    return {
        "phi_solid": float(np.float64(c["solid"]) / np.float64(c["cells"])),
        "interfacial_area_density": float(area / c["cells"]),
        "mean_beam_length": float(4.0 * float(c["pore"]) / float(area)),
        "mean_pore_chord": chords,
        "pore_percolates": c["percolates"],
        "pore_largest_cluster_fraction": float(np.int64(c["largest_cluster"]) / max(1, c["pore"])),
    }


def periodic_chords(pore: np.ndarray, axis: int):
    """(chords in the reference's order, all-pore lines, lines, N) of one axis."""

    rows = axis_lines(pore, axis)
    M, L = rows.shape
    r, s, e = run_table(rows)
    full = rows.all(axis=1)
    wrap = rows[:, 0] & rows[:, -1] & ~full
    at_start = (s == 0) & wrap[r]
    at_end = (e == L) & wrap[r]
    lengths = (e - s).astype(np.int32)
    kept = lengths[~(at_start | at_end)]
    first = np.zeros(M, np.int64)
    last = np.zeros(M, np.int64)
    first[r[at_start]] = lengths[at_start]
    last[r[at_end]] = lengths[at_end]
    joined = (first + last)[wrap].astype(np.int32)
    return np.concatenate([kept, joined]), int(full.sum()), M, L


def chord_histogram(chords: np.ndarray, bins: int) -> np.ndarray:
    h = np.zeros(bins, np.int64)
    c = chords[chords < bins]
    np.add.at(h, c, 1)
    return h


def lineal_path(counts, n_all_pore, n_lines, N):
    c = np.arange(len(counts), dtype=np.int64)
    l = np.arange(1, N + 1, dtype=np.int64)
    segments = (np.maximum(c[None, :] - l[:, None] + 1, 0) * np.asarray(counts)[None, :]).sum(axis=1)
    return (segments + n_all_pore * (l - 1)) / (n_lines * N)


def chord_length_distribution(counts):
    counts = np.asarray(counts)
    total = counts.sum()
    survival = np.cumsum(counts[::-1])[::-1]
    l = np.arange(1, len(counts), dtype=np.int64)
    return counts[1:] / total, survival[1:] / total, l * counts[1:] / np.sum(np.arange(len(counts)) * counts)


def directional_descriptors(phases: np.ndarray) -> list[dict]:
    p = np.asarray(phases)
    pore = p == 0
    N = int(p.shape[0])
    fc = periodic_faces(pore)
    f = fc / fc.sum()
    out = []
    for axis in range(3):
        chords, allp, nlines, n = periodic_chords(pore, axis)
        counts = chord_histogram(chords, N + 1)
        lp = lineal_path(counts, allp, nlines, n)
        vals = chords.astype(float)
        q = np.quantile(vals, list(QUANTILES))
        mean, std = vals.mean(), vals.std()
        below = np.flatnonzero(lp <= np.exp(-1))
        row = {
            "interface_face_count_axis": int(fc[axis]),
            "interface_face_count_total": int(fc.sum()),
            "interface_normal_fraction_axis": float(f[axis]),
            "interface_normal_anisotropy": float(f.max() / f.min()),
            "all_pore_line_count": int(allp),
            "total_axis_lines": int(nlines),
            "all_pore_line_fraction": float(allp / nlines),
            "lineal_integral_cells": float(lp.sum()),
            "lineal_1e_length_cells": float(below[0] + 1) if below.size else float(N),
            "pore_chord_cells_count": int(vals.size),
            "pore_chord_cells_mean": float(mean),
            "pore_chord_cells_std": float(std),
            "pore_chord_cells_max": float(vals.max()),
            "pore_chord_cells_cv": float(std / mean),
        }
        for qq, v in zip(QUANTILES, q):
            row[f"pore_chord_cells_q{int(round(100 * qq))}"] = float(v)
        row["pore_chord_cells_q95_over_mean"] = float(q[3] / mean)
        row["pore_chord_cells_q99_over_mean"] = float(q[4] / mean)
        for l in LINEAL_SAMPLES:
            row[f"lineal_path_L{l:03d}"] = float(lp[l - 1])
        row["lineal_path"] = lp
        row["chord_counts"] = counts
        out.append(row)
    return out


# Small-grid definitions used to pin the functions above (pure Python).

def brute_periodic_chords(pore: np.ndarray, axis: int) -> list[int]:
    p = np.asarray(pore, bool)
    others = [a for a in range(3) if a != axis]
    L = p.shape[axis]
    kept, joined = [], []
    for u in range(p.shape[others[0]]):
        for v in range(p.shape[others[1]]):
            idx = [0, 0, 0]
            idx[others[0]], idx[others[1]] = u, v
            line = []
            for t in range(L):
                idx[axis] = t
                line.append(bool(p[tuple(idx)]))
            runs, t = [], 0
            while t < L:
                if line[t]:
                    t0 = t
                    while t < L and line[t]:
                        t += 1
                    runs.append((t0, t))
                else:
                    t += 1
            if line[0] and line[-1] and not all(line):
                joined.append((runs[0][1] - runs[0][0]) + (runs[-1][1] - runs[-1][0]))
                runs = runs[1:-1]
            kept.extend(b - a for a, b in runs)
    return kept + joined


def brute_clusters(pore: np.ndarray):
    p = np.asarray(pore, bool)
    nx, ny, nz = p.shape
    seen = np.zeros(p.shape, bool)
    sizes, perc = [], [False, False, False]
    for start in zip(*np.nonzero(p)):
        if seen[start]:
            continue
        stack, seen[start], members = [start], True, []
        while stack:
            c = stack.pop()
            members.append(c)
            for di in (-1, 0, 1):
                for dj in (-1, 0, 1):
                    for dk in (-1, 0, 1):
                        d = (c[0] + di, c[1] + dj, c[2] + dk)
                        if 0 <= d[0] < nx and 0 <= d[1] < ny and 0 <= d[2] < nz and p[d] and not seen[d]:
                            seen[d] = True
                            stack.append(d)
        sizes.append(len(members))
        for a, n in enumerate((nx, ny, nz)):
            coords = {m[a] for m in members}
            if 0 in coords and n - 1 in coords:
                perc[a] = True
    return tuple(perc), max(sizes, default=0), len(sizes)
