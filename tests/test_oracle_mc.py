"""The CPU oracle of the multi-bounce Monte Carlo (oracle.mc_properties,
oracle.mc_penetration) against the reference's archived tables and closed forms.

* the archived rows of results/radiative_properties.csv and
  results/penetration_response.csv (and their hit_point_restart/ twins, in
  tests/golden/cases.json): beta to 1e-13 relative, omega to 1e-15 and g to
  1e-13 absolute (the reference sums g with atomics in no fixed order),
  reflected / transmitted / absorbed counts exactly, for rows whose
  microstructure the golden fixtures hold (a few by default, all 216 + 216
  with PORERAD_FULL=1);
* an empty slab: every beam ray crosses it unintercepted (censored at nx) and
  every Lambertian ray is transmitted;
* a solid face at x = 0: every beam ray hits at t = 0 and a reflected ray
  leaves at once, so absorbed / (absorbed + reflected) is eps up to sampling.
"""

from __future__ import annotations

import math
import os

import numpy as np
import pytest

from golden_data import cases, phases_x
from oracle import oracle as O

FULL = os.environ.get("PORERAD_FULL") == "1"
DEFAULT = {("OF_core_phi0.10_s2", "y", 0.6), ("TL_core_phi0.15_s1", "x", 0.3), ("FA_core_phi0.05_s5", "z", 0.9)}


def _seed_mean(values, ses):
    v, s = np.asarray(values, float), np.asarray(ses, float)
    within = float(np.sum(s**2) / v.size**2)
    between = float(v.var(ddof=1) / v.size) if v.size > 1 else 0.0
    return float(v.mean()), float(np.sqrt(max(0.0, within + between)))


def oracle_properties(solid, eps, restart, seeds=(123, 456), n_rays=200_000):
    """beta, omega, g of properties.py:54-329 from the oracle's per-ray output."""

    runs = []
    for seed in seeds:
        o = O.mc_properties(solid, n_rays, seed, eps, restart=restart)
        ev, ce = o["event"][o["event"] >= 0.0], o["censored"][o["censored"] >= 0.0]
        exposure = float(np.sum(ev) + np.sum(ce))
        beta, beta_se = ev.size / exposure, math.sqrt(ev.size) / exposure
        inter, scat = int(o["interactions"].sum()), int(o["scattered"].sum())
        om = scat / inter
        s1, s2 = float(o["mu"].sum()), float(o["mu2"].sum())
        g = s1 / scat
        runs.append((beta, beta_se, om, math.sqrt(om * (1 - om) / inter), g,
                     math.sqrt(max(0.0, (s2 - scat * g * g) / (scat - 1)) / scat)))
    return tuple(_seed_mean([r[2 * k] for r in runs], [r[2 * k + 1] for r in runs])[0] for k in range(3))


def mc_params(table, full=FULL):
    out = []
    for r in cases()[table]:
        key = (r["microstructure"], r["transfer_axis"], r["epsilon"])
        if full or key in DEFAULT:
            out.append(pytest.param(r, id=f"{r['restart']}-{key[0]}-{key[1]}-{key[2]}"))
    return out


@pytest.mark.parametrize("row", mc_params("mc_properties"))
def test_properties_reproduce_archive(row):
    solid = phases_x(row["microstructure"], row["transfer_axis"]) == 1
    beta, omega, g = oracle_properties(solid, row["epsilon"], row["restart"])
    assert abs(beta / row["beta_per_cell"] - 1.0) <= 1e-13
    assert abs(omega - row["omega"]) <= 1e-15
    assert abs(g - row["g"]) <= 1e-13


@pytest.mark.parametrize("row", mc_params("mc_penetration"))
def test_penetration_reproduces_archive(row):
    solid = phases_x(row["microstructure"], row["transfer_axis"]) == 1
    counts = np.zeros(4, np.int64)
    for seed in (321, 654):
        counts += np.bincount(O.mc_penetration(solid, 500_000, seed, row["epsilon"], restart=row["restart"]),
                              minlength=4)
    for k, name in enumerate(("reflectance", "transmittance", "absorptance")):
        assert counts[k] == round(row[name] * row["n_rays"]), name
    assert counts[3] == 0


def test_empty_slab():
    solid = np.zeros((10, 6, 7), np.uint8)
    o = O.mc_properties(solid, 1000, 123, 0.5)
    assert np.all(o["event"] < 0) and np.all(o["censored"] == 10.0)
    assert o["interactions"].sum() == 0
    assert np.array_equal(np.bincount(O.mc_penetration(solid, 2000, 321, 0.5), minlength=4), [0, 2000, 0, 0])


@pytest.mark.parametrize("restart", ["article", "hit_point"])
def test_solid_entry_face(restart):
    solid = np.zeros((10, 6, 7), np.uint8)
    solid[0] = 1
    o = O.mc_properties(solid, 4000, 123, 0.3, restart=restart)
    assert np.all(o["event"] == 0.0) and np.all(o["censored"] < 0)
    assert np.array_equal(o["interactions"], np.ones(4000, np.int64))  # absorbed, or reflected straight out
    fate = np.bincount(O.mc_penetration(solid, 20000, 321, 0.3, restart=restart), minlength=4)
    assert fate[1] == 0 and fate[3] == 0
    assert abs(fate[2] / 20000 - 0.3) <= 4 * math.sqrt(0.3 * 0.7 / 20000)
