"""Morphology descriptors of a two-phase digital microstructure (porerad/descriptors.py).

Phase arrays hold 1 for solid and 0 for pore; lengths are in cells.  The
scalar descriptors (phi_s, A_s, L_m, mean axial pore chords, pore percolation
and the largest 26-connected pore cluster) treat the box faces as ends of the
sample; the directional descriptors (periodic chord-length distribution,
lineal-path function L_a(l), interface-normal fractions f_a) use the periodic
cell.

The integer work (face counts, pore runs, cluster labelling, the chord list of
every axis) runs on the device (``prh_scalar_counts``,
``prh_periodic_chords``); ratios, means and quantiles are formed here from
those integers with the same NumPy operations as the reference, so every value
equals the reference's bit for bit.
"""

from __future__ import annotations

import warnings

import numpy as np

from . import _native

QUANTILES = (0.5, 0.75, 0.9, 0.95, 0.99)
LINEAL_SAMPLES = (1, 2, 4, 8, 16, 32, 64, 128)


def _field(values: np.ndarray) -> "_native.PhaseField":
    return _native.PhaseField(np.ascontiguousarray(values, dtype=np.uint8))


def _pore_field(pore: np.ndarray) -> "_native.PhaseField":
    """A field whose phase-0 cells are the True cells of ``pore``."""

    return _field(~np.asarray(pore, bool))


def _mean_runs(pore_cells: int, runs: int) -> float:
    # np.mean of the run lengths: their sum is the number of pore cells.
    if runs == 0:
        warnings.warn("Mean of empty slice.", RuntimeWarning, stacklevel=3)
        return float("nan")
    return float(np.float64(pore_cells) / np.float64(runs))


# Scalar descriptors

def interface_area(solid: np.ndarray) -> int:
    """Number of cell faces between solid and pore inside the box, A_int (descriptors.py:38-43)."""

    with _field(np.asarray(solid, bool)) as f:
        c = f.scalar_counts(connectivity=False)
    return int(c[_native.DESC_BOX_FACES:_native.DESC_BOX_FACES + 3].sum())


def mean_pore_chords(pore: np.ndarray) -> tuple[float, float, float]:
    """Mean length of the pore runs along x, y and z, runs ending at the box faces (descriptors.py:46-48)."""

    with _pore_field(pore) as f:
        c = f.scalar_counts(connectivity=False)
    return tuple(_mean_runs(int(c[_native.DESC_PORE]), int(c[_native.DESC_PORE_RUNS + a])) for a in range(3))


def _connectivity(c: np.ndarray) -> tuple[tuple[bool, bool, bool], float]:
    bits = int(c[_native.DESC_PERCOLATES])
    largest = int(c[_native.DESC_LARGEST_CLUSTER])
    pore = int(c[_native.DESC_PORE])
    return tuple(bool((bits >> a) & 1) for a in range(3)), float(np.int64(largest) / max(1, pore))


def pore_connectivity(pore: np.ndarray) -> tuple[tuple[bool, bool, bool], float]:
    """Pore percolation along x, y, z and the largest-cluster fraction, 26-connected (descriptors.py:51-60)."""

    with _pore_field(pore) as f:
        return _connectivity(f.scalar_counts(connectivity=True))


def scalar_descriptors(phases: np.ndarray) -> dict:
    """phi_s, A_s = A_int/N^3, L_m = 4 V_pore/A_int, mean axial pore chords and pore connectivity (descriptors.py:63-79)."""

    p = np.asarray(phases)
    with _field(p) as f:
        c = f.scalar_counts(connectivity=True)
    size = int(p.size)
    solid, pore = int(c[_native.DESC_SOLID]), int(c[_native.DESC_PORE])
    area = int(c[_native.DESC_BOX_FACES:_native.DESC_BOX_FACES + 3].sum())
    perc, largest = _connectivity(c)
    return {
        "phi_solid": float(np.float64(solid) / np.float64(size)),
        "interfacial_area_density": float(area / size),
        "mean_beam_length": float(4.0 * float(pore) / float(area)),
        "mean_pore_chord": tuple(_mean_runs(pore, int(c[_native.DESC_PORE_RUNS + a])) for a in range(3)),
        "pore_percolates": perc,
        "pore_largest_cluster_fraction": largest,
    }


# Directional descriptors

def periodic_chords(pore: np.ndarray, axis: int):
    """Pore chords along one axis of the periodic cell (descriptors.py:82-109).

    Runs touching both ends of a line are merged into one chord; a line lying
    entirely in the pore gives one chord of length N.  Returns the chord
    lengths (int32, in the reference's order), the number of all-pore lines,
    the number of lines and N.
    """

    pore = np.asarray(pore, bool)
    with _pore_field(pore) as f:
        chords, allp, _ = f.periodic_chords(int(axis), 1)
    n = int(pore.shape[axis])
    return chords, allp, int(pore.size) // n, n


def lineal_path(counts: np.ndarray, n_all_pore: int, n_lines: int, N: int) -> np.ndarray:
    """L_a(l), l = 1..N, from the chord-count histogram of the periodic cell (descriptors.py:112-117).

    A chord of length c holds max(c - l + 1, 0) segments of length l; an
    all-pore line adds l - 1 segments that wrap across its ends.
    """

    counts = np.asarray(counts)
    c = np.arange(len(counts), dtype=np.int64)
    l = np.arange(1, int(N) + 1, dtype=np.int64)
    segments = (np.maximum(c[None, :] - l[:, None] + 1, 0) * counts[None, :]).sum(axis=1)
    return (segments + n_all_pore * (l - 1)) / (n_lines * N)


def interface_face_counts(solid: np.ndarray) -> np.ndarray:
    """Solid-pore faces of the periodic cell whose normal is parallel to x, y and z (descriptors.py:120-123)."""

    with _field(~np.asarray(solid, bool)) as f:
        c = f.scalar_counts(connectivity=False)
    return c[_native.DESC_PERIODIC_FACES:_native.DESC_PERIODIC_FACES + 3].astype(np.int64)


def chord_length_distribution(counts: np.ndarray):
    """Chord probability, survival probability and pore-length fraction for l = 1..N (descriptors.py:126-132)."""

    counts = np.asarray(counts)
    total = counts.sum()
    length_sum = np.sum(np.arange(len(counts)) * counts)
    survival = np.cumsum(counts[::-1])[::-1]
    l = np.arange(1, len(counts), dtype=np.int64)
    return counts[1:] / total, survival[1:] / total, l * counts[1:] / length_sum


def directional_descriptors(phases: np.ndarray) -> list[dict]:
    """Per-axis directional descriptors for x, y, z (descriptors.py:135-180).

    Returns three dicts with the summary values, the lineal-path function
    ``lineal_path`` (l = 1..N) and the chord-count histogram ``chord_counts``
    (index = chord length in cells).  One upload serves the three axes.
    """

    p = np.asarray(phases)
    # The field of the directional descriptors is pore = (phase == 0), solid = the rest.
    with _field(p != 0) as fld:
        c = fld.scalar_counts(connectivity=False)
        fc = c[_native.DESC_PERIODIC_FACES:_native.DESC_PERIODIC_FACES + 3].astype(np.int64)
        N = int(p.shape[0])
        per_axis = [fld.periodic_chords(a, N + 1) for a in range(3)]
    f = fc / fc.sum()
    out = []
    for axis, (chords, allp, counts) in enumerate(per_axis):
        n = int(p.shape[axis])
        nlines = int(p.size) // n
        lp = lineal_path(counts, allp, nlines, n)
        vals = chords.astype(float)
        q = np.quantile(vals, list(QUANTILES))
        below = lp <= np.exp(-1)
        mean = vals.mean()
        std = vals.std()
        row = {
            "interface_face_count_axis": int(fc[axis]),
            "interface_face_count_total": int(fc.sum()),
            "interface_normal_fraction_axis": float(f[axis]),
            "interface_normal_anisotropy": float(f.max() / f.min()),
            "all_pore_line_count": int(allp),
            "total_axis_lines": int(nlines),
            "all_pore_line_fraction": float(allp / nlines),
            "lineal_integral_cells": float(lp.sum()),
            "lineal_1e_length_cells": float(np.flatnonzero(below)[0] + 1) if below.any() else float(N),
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
