"""Pinning of the descriptor oracle (CPU).

The oracle restates porerad/descriptors.py; these checks tie it to the
reference's own data and to definitions evaluated independently:
  * data/microstructures.csv (phi_s, A_s, L_m and the mean axial pore chords
    of every microstructure; exact) for every fixture geometry;
  * the acceptance rules the core library was filtered with
    (porerad/geometry.py:261-281), which fix percolation, the largest pore
    cluster and the chord ratios of every accepted core geometry;
  * pure-Python line walks and a 26-neighbour breadth-first search on small
    grids (run order, joined end runs, face counts, clusters);
  * closed forms: L(1) = porosity, the lineal path of one chord, parallel
    plates.
"""

from __future__ import annotations

import math
import os

import numpy as np
import pytest

from golden_data import GOLDEN, cases, phases
from oracle import descriptors_oracle as D

TABLE = {r["id"]: r for r in cases()["microstructures"]}


def fixture_names():
    with np.load(os.path.join(GOLDEN, "geometries.npz"), allow_pickle=False) as z:
        return sorted({k.split("__")[0] for k in z.files})


SMALL_SHAPES = [(5, 4, 6), (1, 3, 7), (6, 6, 6), (7, 1, 5), (3, 8, 1), (9, 5, 4)]


def small_field(shape, porosity, seed, three_phase=False):
    rng = np.random.default_rng(seed)
    p = (rng.random(shape) >= porosity).astype(np.uint8)
    if three_phase:
        p[rng.random(shape) < 0.1] = 2
    return p


@pytest.mark.parametrize("name", fixture_names())
def test_scalar_descriptors_match_microstructure_table(name):
    row = TABLE[name]
    d = D.scalar_descriptors(phases(name))
    assert d["phi_solid"] == row["realized_phi"]
    assert d["interfacial_area_density"] == row["interfacial_area_density_per_cell"]
    assert d["mean_beam_length"] == row["mean_beam_length_cells"]
    assert d["mean_pore_chord"] == (row["mean_pore_chord_x_cells"], row["mean_pore_chord_y_cells"],
                                    row["mean_pore_chord_z_cells"])


@pytest.mark.parametrize("name", [n for n in fixture_names() if "_core_" in n])
def test_accepted_core_geometries_satisfy_the_acceptance_rule(name):
    row = TABLE[name]
    d = D.scalar_descriptors(phases(name))
    assert abs(d["phi_solid"] - row["target_phi"]) <= 3e-3
    c = d["mean_pore_chord"]
    s = row["structure"]
    if s == "OF":
        assert all(d["pore_percolates"]) and d["pore_largest_cluster_fraction"] >= 0.99
    elif s == "TL":
        assert max(c) / min(c) >= 2.0
    elif s == "FA":
        assert c[0] / (0.5 * (c[1] + c[2])) >= 2.0
    else:
        assert max(c) / min(c) < 2.0


def brute_counts(p):
    nx, ny, nz = p.shape
    solid, pore = p == 1, p == 0
    box = [0, 0, 0]
    per = [0, 0, 0]
    runs = [0, 0, 0]
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                c = (i, j, k)
                for a in range(3):
                    nxt = list(c)
                    nxt[a] += 1
                    if nxt[a] < p.shape[a] and solid[c] != solid[tuple(nxt)]:
                        box[a] += 1
                    nxt[a] %= p.shape[a]
                    if pore[c] != pore[tuple(nxt)]:
                        per[a] += 1
                    prv = list(c)
                    prv[a] -= 1
                    if pore[c] and (prv[a] < 0 or not pore[tuple(prv)]):
                        runs[a] += 1
    return box, per, runs


@pytest.mark.parametrize("shape", SMALL_SHAPES)
@pytest.mark.parametrize("porosity", [0.2, 0.6, 0.9])
def test_counts_against_cell_walks(shape, porosity):
    p = small_field(shape, porosity, seed=sum(shape) + int(10 * porosity), three_phase=True)
    c = D.scalar_counts(p)
    box, per, runs = brute_counts(p)
    assert c["box_faces"].tolist() == box
    assert c["periodic_faces"].tolist() == per
    assert c["pore_runs"].tolist() == runs
    assert c["solid"] == int((p == 1).sum()) and c["pore"] == int((p == 0).sum())
    perc, largest, count = D.brute_clusters(p == 0)
    assert c["percolates"] == perc and c["largest_cluster"] == largest and c["clusters"] == count


@pytest.mark.parametrize("shape", SMALL_SHAPES + [(8, 8, 8)])
@pytest.mark.parametrize("porosity", [0.3, 0.7, 0.95])
def test_periodic_chords_against_line_walks(shape, porosity):
    pore = small_field(shape, porosity, seed=7 * sum(shape) + int(100 * porosity)) == 0
    for axis in range(3):
        chords, allp, lines, n = D.periodic_chords(pore, axis)
        assert chords.tolist() == D.brute_periodic_chords(pore, axis)
        assert lines * n == pore.size and n == pore.shape[axis]
        assert allp == int(np.all(pore, axis=axis).sum())
        # every pore cell lies on exactly one chord
        assert int(chords.sum()) == int(pore.sum())


def test_lineal_path_closed_forms():
    rng = np.random.default_rng(3)
    pore = rng.random((10, 10, 10)) < 0.7
    for axis in range(3):
        chords, allp, lines, n = D.periodic_chords(pore, axis)
        counts = D.chord_histogram(chords, 11)
        lp = D.lineal_path(counts, allp, lines, n)
        assert lp[0] == pytest.approx(pore.mean(), rel=0, abs=1e-15)
        assert np.all(np.diff(lp) <= 0)
    # one chord of length 4 on one of 6 lines of length 8
    counts = np.zeros(9, np.int64)
    counts[4] = 1
    lp = D.lineal_path(counts, 0, 6, 8)
    assert lp.tolist() == [(4 - l + 1) / 48 if l <= 4 else 0.0 for l in range(1, 9)]
    # an all-pore line of length 8 contains 8 segments of every length
    counts = np.zeros(9, np.int64)
    counts[8] = 1
    assert np.all(D.lineal_path(counts, 1, 1, 8) == 1.0)


def test_parallel_plates():
    N, gap = 128, 3
    planes = N // (gap + 1)
    p = np.zeros((N, N, N), np.uint8)
    p[:: gap + 1] = 1  # solid planes x = 0, 4, 8, ...
    with np.errstate(divide="ignore"):
        rows = D.directional_descriptors(p)  # no face normal to y or z: anisotropy 1/0
    x = rows[0]
    assert x["interface_normal_anisotropy"] == math.inf
    assert x["pore_chord_cells_count"] == planes * N * N
    assert x["pore_chord_cells_mean"] == gap and x["pore_chord_cells_std"] == 0.0
    assert all(x[f"pore_chord_cells_q{q}"] == gap for q in (50, 75, 90, 95, 99))
    assert x["interface_face_count_axis"] == 2 * planes * N * N
    assert x["interface_face_count_total"] == 2 * planes * N * N
    assert x["all_pore_line_count"] == 0 and x["total_axis_lines"] == N * N
    assert x["lineal_path_L001"] == 0.75 and x["lineal_path_L004"] == 0.0
    assert x["lineal_path_L002"] == 0.5 and x["lineal_1e_length_cells"] == 3.0
    for r in rows[1:]:
        # lines inside the pore slabs are all pore; lines inside the plates hold no chord
        assert r["all_pore_line_count"] == gap * planes * N
        assert r["pore_chord_cells_count"] == gap * planes * N
        assert r["pore_chord_cells_mean"] == N and r["lineal_path_L001"] == 0.75
        assert r["lineal_path_L128"] == 0.75
    d = D.scalar_descriptors(p)
    assert d["mean_pore_chord"] == (gap, N, N) and d["pore_percolates"] == (False, True, True)
    assert d["pore_largest_cluster_fraction"] == 1 / planes
    # interfaces inside the box: one beside the plane x = 0, two beside every other plane
    assert d["mean_beam_length"] == 4 * 0.75 * N**3 / ((2 * planes - 1) * N * N)
