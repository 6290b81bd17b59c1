"""Pinning of the microstructure-generator oracle (CPU).

  * The phase SHA-256 of data/microstructures.csv: a representative row of
    every structure and set by default, all 66 rows with PORERAD_FULL=1.
  * The C stamping's membership test against NumPy's own evaluation of the
    same shapes (R.T @ X through NumPy's BLAS, float remainder, powers),
    which fixes the FMA-chain reading of the BLAS product.
"""

from __future__ import annotations

import numpy as np
import pytest

from golden_data import cases
from oracle import geometry_oracle as G

TABLE = {r["id"]: r for r in cases()["microstructures"]}
REPRESENTATIVE = ["OF_core_phi0.05_s3", "FI_core_phi0.10_s3", "FA_core_phi0.15_s2", "TL_core_phi0.05_s3",
                  "FA_canonical_phi0.10_N160", "FI_canonical_phi0.10_N192", "TL_canonical_phi0.10_N160",
                  "TL_canonical_phi0.10_N192", "FA_matched_pair3", "FI_matched_pair1"]


@pytest.mark.parametrize("name", REPRESENTATIVE)
def test_phase_sha256(name):
    assert G.sha256(G.generate(TABLE[name])) == TABLE[name]["phase_sha256"]


@pytest.mark.slow
@pytest.mark.parametrize("name", sorted(set(TABLE) - set(REPRESENTATIVE)))
def test_phase_sha256_all_rows(name):
    assert G.sha256(G.generate(TABLE[name])) == TABLE[name]["phase_sha256"]


def numpy_shape(N, center, radii, kind, R):
    """The membership test written with NumPy array operations."""

    half = float(N) / 2.0
    ax = [(np.arange(N) - c + half) % float(N) - half for c in center]
    x3, y3, z3 = np.meshgrid(*ax, indexing="ij")
    local = R.T @ np.stack([x3.ravel(), y3.ravel(), z3.ravel()], axis=0)
    q = [(local[i] / radii[i]) ** 2 for i in range(3)]
    if kind == G.ELLIPSOID:
        inside = q[0] + q[1] + q[2] <= 1.0
    else:
        inside = (q[0] + q[1] <= 1.0) & (np.abs(local[2]) <= radii[2])
    return inside.reshape((N, N, N)).astype(np.uint8)


@pytest.mark.parametrize("seed", range(6))
def test_stamping_matches_numpy(seed):
    rng = np.random.default_rng(seed)
    N = int(rng.integers(20, 48))
    for kind in (G.ELLIPSOID, G.CYLINDER):
        v = rng.normal(size=3)
        R = G.frame(v / np.linalg.norm(v), G.E_Z)
        center = tuple(rng.random(3) * N)
        radii = (float(rng.uniform(1, 6)), float(rng.uniform(1, 6)), float(rng.uniform(2, 0.5 * N)))
        canvas = G.Canvas(N)
        canvas.add(center, radii, kind, R)
        assert np.array_equal(canvas.finish(), numpy_shape(N, center, radii, kind, R))


def test_long_cylinder_roll():
    """An integer-centred long cylinder equals the one built at the origin, rolled by whole cells."""

    N = 24
    v = np.array([0.3, -0.5, 0.8])
    R = G.frame(v / np.linalg.norm(v), G.E_Z)
    a = G.Canvas(N)
    a.add((5.0, 7.0, 11.0), (2.2, 2.2, 4.5 * N), G.CYLINDER, R)
    b = G.Canvas(N)
    b.add((0.0, 0.0, 0.0), (2.2, 2.2, 4.5 * N), G.CYLINDER, R)
    assert np.array_equal(a.finish(), np.roll(b.finish(), shift=(5, 7, 11), axis=(0, 1, 2)))
