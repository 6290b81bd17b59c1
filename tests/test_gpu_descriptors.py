"""Device morphology descriptors against the oracle (bit-exact) and the microstructure table."""

from __future__ import annotations

import math
import os

import numpy as np
import pytest

from golden_data import GOLDEN, cases, phases
from oracle import descriptors_oracle as D

pytestmark = pytest.mark.gpu

TABLE = {r["id"]: r for r in cases()["microstructures"]}


def fixture_names():
    with np.load(os.path.join(GOLDEN, "geometries.npz"), allow_pickle=False) as z:
        return sorted({k.split("__")[0] for k in z.files})


@pytest.fixture(scope="module")
def desc():
    from porerad import descriptors

    return descriptors


@pytest.fixture(scope="module")
def native():
    from porerad import _native

    return _native


def assert_rows_equal(got, want):
    assert got.keys() == want.keys()
    for k in want:
        if isinstance(want[k], np.ndarray):
            assert np.array_equal(got[k], want[k]), k
        else:
            assert got[k] == want[k] or (math.isnan(got[k]) and math.isnan(want[k])), k


@pytest.mark.parametrize("name", fixture_names())
def test_fixture_geometries(desc, name):
    p = phases(name)
    row = TABLE[name]
    s = desc.scalar_descriptors(p)
    assert s == D.scalar_descriptors(p)
    assert s["phi_solid"] == row["realized_phi"]
    assert s["interfacial_area_density"] == row["interfacial_area_density_per_cell"]
    assert s["mean_beam_length"] == row["mean_beam_length_cells"]
    assert s["mean_pore_chord"] == (row["mean_pore_chord_x_cells"], row["mean_pore_chord_y_cells"],
                                    row["mean_pore_chord_z_cells"])
    got, want = desc.directional_descriptors(p), D.directional_descriptors(p)
    for g, w in zip(got, want):
        assert_rows_equal(g, w)


# (4100, 2, 3): a histogram longer than the kernel's LDS copy (4096 bins)
SHAPES = [(5, 4, 6), (1, 3, 7), (7, 1, 5), (3, 8, 1), (1, 1, 1), (1, 1, 9), (9, 1, 1), (33, 17, 65), (64, 64, 64),
          (4100, 2, 3), (130, 70, 129)]


def field(shape, porosity, seed):
    rng = np.random.default_rng(seed)
    p = (rng.random(shape) >= porosity).astype(np.uint8)
    p[rng.random(shape) < 0.05] = 2  # a third phase: neither solid nor pore
    return p


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("porosity", [0.25, 0.7, 1.0])
def test_counts_and_chords(native, shape, porosity):
    p = field(shape, porosity, seed=sum(shape) + int(100 * porosity))
    if porosity == 1.0:
        p[:] = 0
    want = D.scalar_counts(p)
    with native.PhaseField(p) as f:
        c = f.scalar_counts(connectivity=True)
        assert c[native.DESC_CELLS] == p.size
        assert c[native.DESC_SOLID] == want["solid"] and c[native.DESC_PORE] == want["pore"]
        assert c[native.DESC_BOX_FACES:native.DESC_BOX_FACES + 3].tolist() == want["box_faces"].tolist()
        assert c[native.DESC_PORE_RUNS:native.DESC_PORE_RUNS + 3].tolist() == want["pore_runs"].tolist()
        assert c[native.DESC_PERIODIC_FACES:native.DESC_PERIODIC_FACES + 3].tolist() == want["periodic_faces"].tolist()
        assert c[native.DESC_LARGEST_CLUSTER] == want["largest_cluster"]
        assert c[native.DESC_CLUSTERS] == want["clusters"]
        assert [bool((int(c[native.DESC_PERCOLATES]) >> a) & 1) for a in range(3)] == list(want["percolates"])
        assert np.array_equal(f.read(), p)
        for axis in range(3):
            chords, allp, hist = f.periodic_chords(axis, p.shape[0] + 1)
            w_chords, w_allp, _, _ = D.periodic_chords(p == 0, axis)
            assert chords.dtype == np.int32 and np.array_equal(chords, w_chords)
            assert allp == w_allp
            assert np.array_equal(hist, D.chord_histogram(w_chords, p.shape[0] + 1))


def test_percolation_threshold_clusters(native):
    """Many clusters of every size near the site-percolation threshold of the 26-neighbour lattice."""

    p = field((96, 96, 96), 0.1, seed=11)
    p[p == 2] = 1
    want = D.scalar_counts(p)
    assert want["clusters"] > 1000
    with native.PhaseField(p) as f:
        c = f.scalar_counts(connectivity=True)
    assert c[native.DESC_CLUSTERS] == want["clusters"]
    assert c[native.DESC_LARGEST_CLUSTER] == want["largest_cluster"]


def test_api_functions(desc):
    p = phases("OF_core_phi0.10_s2")
    pore, solid = p == 0, p == 1
    assert desc.interface_area(solid) == int(D.box_faces(solid).sum())
    assert desc.mean_pore_chords(pore) == D.scalar_descriptors(p)["mean_pore_chord"]
    perc, largest, _ = D.clusters(pore)
    assert desc.pore_connectivity(pore) == (perc, float(np.int64(largest) / int(pore.sum())))
    assert np.array_equal(desc.interface_face_counts(solid), D.periodic_faces(pore))
    for axis in range(3):
        got = desc.periodic_chords(pore, axis)
        want = D.periodic_chords(pore, axis)
        assert np.array_equal(got[0], want[0]) and got[1:] == want[1:]
    counts = D.directional_descriptors(p)[0]["chord_counts"]
    for g, w in zip(desc.chord_length_distribution(counts), D.chord_length_distribution(counts)):
        assert np.array_equal(g, w)


def test_edge_cases(desc):
    solid = np.ones((4, 5, 6), np.uint8)
    with pytest.raises(ZeroDivisionError):
        desc.scalar_descriptors(solid)  # no interface: L_m = 4 V_pore / 0, as in the reference
    with pytest.warns(RuntimeWarning):
        assert all(math.isnan(v) for v in desc.mean_pore_chords(solid == 0))
    assert desc.pore_connectivity(solid == 0) == ((False, False, False), 0.0)
    chords, allp, lines, n = desc.periodic_chords(solid == 0, 1)
    assert chords.size == 0 and allp == 0 and lines == 24 and n == 5
    pore = np.zeros((4, 5, 6), np.uint8)
    assert desc.pore_connectivity(pore == 0) == ((True, True, True), 1.0)
    chords, allp, lines, n = desc.periodic_chords(pore == 0, 2)
    assert chords.tolist() == [6] * 20 and allp == 20
