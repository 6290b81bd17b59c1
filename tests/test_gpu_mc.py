"""The device multi-bounce Monte Carlo (prh_mc_properties, prh_mc_penetration)
through porerad.properties / porerad.penetration and the C ABI.

* against the CPU oracle on the same seeds: per-ray first path lengths
  bit-identical, interaction and reflection counts equal, the sums of mu and
  mu^2 within 1e-12 relative (summed in another order), penetration counts
  equal -- for both restart rules;
* against the archived tables (tests/golden/cases.json, from
  results/radiative_properties.csv, results/penetration_response.csv and
  their hit_point_restart/ twins): beta to 1e-13 relative, omega to 1e-15, g
  to 1e-13, reflected / transmitted / absorbed counts exactly; every row of
  two microstructures by default, all 432 with PORERAD_FULL=1;
* closed forms: empty slab, solid entry face; a bad restart rule raises.
"""

from __future__ import annotations

import os

import numpy as np
import pytest

from golden_data import phases, phases_x
from oracle import oracle as O
from test_oracle_mc import DEFAULT, mc_params

pytestmark = pytest.mark.gpu

GPU_DEFAULT = {"OF_core_phi0.10_s2", "TL_core_phi0.15_s1"}


@pytest.fixture(scope="module")
def mc():
    from porerad import _native, penetration, properties

    _native.lib()
    return _native, properties, penetration


def _device_properties(mc, solid, n, seed, eps, restart):
    _native = mc[0]
    ev, ce = np.empty(n), np.empty(n)
    counters, sums = np.zeros(2, np.int64), np.zeros(2)
    geo = _native.Geometry(solid, 1.0)
    try:
        _native.check(_native.lib().prh_mc_properties(geo.handle, n, seed, eps, 1000, int(restart == "hit_point"),
                                                      _native.ptr(ev), _native.ptr(ce), _native.ptr(counters),
                                                      _native.ptr(sums)), "prh_mc_properties")
    finally:
        geo.close()
    return ev, ce, counters, sums


@pytest.mark.parametrize("restart", ["article", "hit_point"])
def test_rays_match_oracle(mc, restart):
    _native, properties, penetration = mc
    solid = (phases_x("FI_core_phi0.10_s3", "y") == 1).astype(np.uint8)
    ev, ce, counters, sums = _device_properties(mc, solid, 50_000, 777, 0.45, restart)
    o = O.mc_properties(solid, 50_000, 777, 0.45, restart=restart)
    assert np.array_equal(ev, o["event"]) and np.array_equal(ce, o["censored"])
    assert counters[0] == o["interactions"].sum() and counters[1] == o["scattered"].sum()
    assert abs(sums[0] - o["mu"].sum()) <= 1e-12 * abs(o["mu"].sum())
    assert abs(sums[1] - o["mu2"].sum()) <= 1e-12 * abs(o["mu2"].sum())
    got = penetration.trace_penetration(phases_x("FI_core_phi0.10_s3", "y"), "x", 0.45, n_rays=80_000, seed=99,
                                        restart=restart)
    fate = np.bincount(O.mc_penetration(solid, 80_000, 99, 0.45, restart=restart), minlength=4)
    for k, name in enumerate(("reflectance", "transmittance", "absorptance")):
        assert got[name] == fate[k] / 80_000, name
    assert got["terminated"] == fate[3]


def _rows(table):
    full = mc_params(table, full=True)
    if os.environ.get("PORERAD_FULL") == "1":
        return full
    return [p for p in full if p.values[0]["microstructure"] in GPU_DEFAULT or
            (p.values[0]["microstructure"], p.values[0]["transfer_axis"], p.values[0]["epsilon"]) in DEFAULT]


@pytest.mark.parametrize("row", _rows("mc_properties"))
def test_properties_reproduce_archive(mc, row):
    _, properties, _ = mc
    got = properties.identify_properties(phases(row["microstructure"]), row["transfer_axis"], row["epsilon"],
                                         restart=row["restart"])
    assert got["n_rays"] == row["n_rays"]
    assert abs(got["beta"] / row["beta_per_cell"] - 1.0) <= 1e-13
    assert abs(got["omega"] - row["omega"]) <= 1e-15
    assert abs(got["g"] - row["g"]) <= 1e-13
    assert abs(got["beta_se"] / row["beta_se_per_cell"] - 1.0) <= 1e-9


@pytest.mark.parametrize("row", _rows("mc_penetration"))
def test_penetration_reproduces_archive(mc, row):
    _, _, penetration = mc
    got = penetration.penetration_response(phases(row["microstructure"]), row["transfer_axis"], row["epsilon"],
                                           restart=row["restart"])
    assert got["n_rays"] == row["n_rays"] and got["terminated"] == 0
    for name in ("reflectance", "transmittance", "absorptance"):
        assert round(got[name] * row["n_rays"]) == round(row[name] * row["n_rays"]), name


def test_closed_forms(mc):
    _native, properties, penetration = mc
    empty = np.zeros((10, 6, 7), np.uint8)
    ev, ce, counters, _ = _device_properties(mc, empty, 1000, 123, 0.5, "article")
    assert np.all(ev < 0) and np.all(ce == 10.0) and counters[0] == 0
    r = penetration.trace_penetration(empty, "x", 0.5, n_rays=2000, seed=321)
    assert r["transmittance"] == 1.0 and r["reflectance"] == 0.0
    wall = np.zeros((10, 6, 7), np.uint8)
    wall[0] = 1
    ev, ce, counters, _ = _device_properties(mc, wall, 4000, 123, 0.3, "hit_point")
    assert np.all(ev == 0.0) and counters[0] == 4000
    with pytest.raises(ValueError, match="restart must be one of"):
        properties.trace_properties(wall, "x", 0.3, restart="nowhere")
