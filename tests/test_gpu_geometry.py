"""Device microstructure generators: every microstructure of the table by phase SHA-256, and the oracle."""

from __future__ import annotations

import hashlib

import numpy as np
import pytest

from golden_data import cases
from oracle import geometry_oracle as G

pytestmark = pytest.mark.gpu

TABLE = {r["id"]: r for r in cases()["microstructures"]}


@pytest.fixture(scope="module")
def geo():
    from porerad import geometry

    return geometry


def generate(geo, row):
    s, kind, N = row["structure"], row["set"], int(row["N"])
    if kind == "core":
        return geo.build(s, N, row["target_phi"], int(row["seed"]))
    if kind == "canonical":
        return geo.canonical(s, N)
    return geo.matched(s, int(row["id"].rsplit("pair", 1)[1]), N)


def sha(p):
    return hashlib.sha256(np.ascontiguousarray(p, np.uint8).tobytes()).hexdigest()


@pytest.mark.parametrize("name", sorted(TABLE))
def test_table_microstructure(geo, name):
    p = generate(geo, TABLE[name])
    assert p.dtype == np.uint8 and p.shape == (TABLE[name]["N"],) * 3
    assert sha(p) == TABLE[name]["phase_sha256"]


@pytest.mark.parametrize("seed", range(4))
def test_shape_mask_against_oracle(geo, seed):
    rng = np.random.default_rng(100 + seed)
    N = int(rng.integers(17, 61))
    for kind, k in (("ellipsoid", G.ELLIPSOID), ("cylinder", G.CYLINDER)):
        v = rng.normal(size=3)
        R = G.frame(v / np.linalg.norm(v), G.E_Z)
        center = tuple(rng.random(3) * N)
        # radii up to beyond half the cell: the reach covers every cell
        radii = (float(rng.uniform(0.6, 0.7 * N)), float(rng.uniform(0.6, 8)), float(rng.uniform(1, 0.5 * N)))
        canvas = G.Canvas(N)
        canvas.add(center, radii, k, R)
        assert np.array_equal(geo.shape_mask(N, center, radii, kind, R), canvas.finish().astype(bool))


def test_shape_mask_general_matrix(geo):
    """A matrix that is not a rotation: every cell is tested (no reach shortcut)."""

    N = 29
    R = np.array([[2.0, 0.3, 0.0], [0.1, 0.5, 0.2], [0.0, -0.4, 1.5]])
    for kind, k in (("ellipsoid", G.ELLIPSOID), ("cylinder", G.CYLINDER)):
        canvas = G.Canvas(N)
        canvas.add((3.3, 20.1, 11.7), (4.0, 2.5, 6.0), k, R)
        assert np.array_equal(geo.shape_mask(N, (3.3, 20.1, 11.7), (4.0, 2.5, 6.0), kind, R),
                              canvas.finish().astype(bool))


def test_threshold_any_shape(geo):
    field = np.random.default_rng(4).normal(size=(7, 300))
    q = np.quantile(field.ravel(), 1.0 - 0.3)
    assert np.array_equal(geo.threshold(field, 0.3), (field >= q).astype(np.uint8))


def test_long_cylinders_against_oracle(geo):
    N = 40
    v = np.array([0.2, 0.9, -0.4])
    R = G.frame(v / np.linalg.norm(v), G.E_Z)
    for center in ((3.0, 17.0, 39.0), (3.4, 17.0, 39.0), (0.0, 0.0, 0.0)):
        canvas = G.Canvas(N)
        canvas.add(center, (2.6, 2.6, 4.5 * N), G.CYLINDER, R)
        assert np.array_equal(geo.shape_mask(N, center, (2.6, 2.6, 4.5 * N), "cylinder", R),
                              canvas.finish().astype(bool))
    assert np.array_equal(geo._long_cylinder(N, 1.7, 5.0 * N, R), geo.shape_mask(N, (0.0, 0.0, 0.0),
                                                                                (1.7, 1.7, 5.0 * N), "cylinder", R))


def test_small_foams_and_fibers_against_oracle(geo):
    for N, seed in ((37, 3), (50, 8)):
        points, edges = geo.foam_graph(N, seed, 9)
        assert np.array_equal(geo.render_foam(N, points, edges, 1.3, 1.56), G.render_foam(N, points, edges, 1.3, 1.56))
        centers, rotations = geo.fiber_lines(N, seed, 5, aligned=seed % 2 == 0, continuous_centers=seed > 5)
        assert np.array_equal(geo.render_fibers(N, centers, rotations, 1.9, 9.0),
                              G.render_fibers(N, centers, rotations, 1.9, 9.0))


def test_bisect_radius_api(geo):
    N = 48
    points, edges = geo.foam_graph(N, 5, 12)
    r_host, p_host = geo._bisect_radius(lambda r: geo.render_foam(N, points, edges, r, 1.2 * r), 0.1, 0.2, 4.0)
    r_ora, p_ora = G.bisect_radius(lambda r: G.render_foam(N, points, edges, r, 1.2 * r), 0.1, 0.2, 4.0)
    assert r_host == r_ora and np.array_equal(p_host, p_ora)


@pytest.mark.parametrize("q", [0.0, 0.37, 0.5, 0.9, 1.0])
def test_threshold_quantile(geo, q):
    rng = np.random.default_rng(int(100 * q))
    field = rng.normal(size=(30, 30, 30))
    field[rng.random(field.shape) < 0.2] = 0.25  # ties at the threshold
    got = geo.threshold(field, 1.0 - q)
    assert np.array_equal(got, G.threshold(field, 1.0 - q))


def test_threshold_with_noise(geo):
    field, noise = geo._lamellar_terms(40, 9, 5, 1.7)
    full = geo.lamellar_field(40, 9, 5, 1.7)
    assert np.array_equal(full, G.lamellar_field(40, 9, 5, 1.7))
    assert np.array_equal(geo._threshold_layers(40, 9, 5, 1.7, 0.12), G.threshold(full, 0.12))


def test_core_realizations(geo):
    """The first accepted TL seeds at phi_s = 0.05 are the table's, with its phases."""

    got = geo.core_realizations("TL", 0.05, n_accept=2)
    rows = sorted((r for r in TABLE.values() if r["structure"] == "TL" and r["set"] == "core"
                   and r["target_phi"] == 0.05), key=lambda r: r["seed"])[:2]
    assert [s for s, _ in got] == [r["seed"] for r in rows]
    assert [sha(p) for _, p in got] == [r["phase_sha256"] for r in rows]
