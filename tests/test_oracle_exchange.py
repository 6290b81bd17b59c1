"""The FP64 exchange oracle reproduces the reference's archived radiosity results.

* results/prescribed_field.csv: q_h, q_c, q_i of every condition (T_m, eps) of
  a prescribed-field case, recomputed from the case's seeded ray sets
  (R = 32 per block) through the oracle tracer, operator and series.
* Closed form: an empty box is two parallel gray plates,
  q = sigma (T_h^4 - T_c^4) / (2 / eps - 1).
* An isothermal enclosure has zero net power everywhere (closure + symmetry).
"""

import numpy as np
import pytest

from golden_data import cases, phases_x
from oracle import exchange_oracle as X
from oracle import oracle as O


def _pf(case_id):
    return next(c for c in cases()["prescribed_field"] if c["case_id"] == case_id)


def _operator(case):
    P = O.build_surface_patches(phases_x(case["microstructure"], case["axis"]), case["H_m"])
    R = case["rays_per_patch_per_block"]
    tables = [O.first_hit_counts(P, R, s) for s in case["final_seeds"]]
    o, r, c = O.pool_counts(tables, P.n_patches)
    rays = np.full(P.n_patches, R * len(case["final_seeds"]), np.int64)
    return P, X.OracleOperator(o, r, c, rays, P.patch_area_m2)


@pytest.mark.parametrize("case_id,rows", [
    ("pf_OF_core_phi0.05_s3_x", [0, 4, 8]),
    pytest.param("pf_TL_core_phi0.10_s0_y", list(range(9)), marks=pytest.mark.slow),
    pytest.param("pf_FA_core_phi0.10_s2_z", list(range(9)), marks=pytest.mark.slow),
    pytest.param("pf_FI_core_phi0.05_s3_y", list(range(9)), marks=pytest.mark.slow),
])
def test_prescribed_field_channels_match_archive(case_id, rows):
    case = _pf(case_id)
    P, op = _operator(case)
    assert op.row_sum_error <= 1e-10
    for k in rows:
        ref = case["results"][k]
        q_h, q_c, q_i, closure, *_ = X.prescribed_channels(op, P, ref["T_mean_K"], ref["delta_T_K"], ref["epsilon"])
        scale = max(abs(ref["q_h_W_m2"]), abs(ref["q_c_W_m2"]))
        assert abs(q_h - ref["q_h_W_m2"]) <= 1e-10 * scale
        assert abs(q_c - ref["q_c_W_m2"]) <= 1e-10 * scale
        assert abs(q_i - ref["q_i_W_m2"]) <= 1e-10 * scale
        assert closure <= 1e-4


@pytest.mark.parametrize("eps", [0.3, 0.6, 0.9])
def test_empty_box_is_parallel_plates(eps):
    P = O.build_surface_patches(np.zeros((6, 5, 5), np.uint8), 0.006)
    o, r, c = O.first_hit_counts(P, 40, 11)
    op = X.OracleOperator(o, r, c, np.full(P.n_patches, 40), P.patch_area_m2)
    T_m, dT = 500.0, 400.0
    q_h, q_c, q_i, *_ = X.prescribed_channels(op, P, T_m, dT, eps)
    Th, Tc = T_m + dT / 2, T_m - dT / 2
    exact = X.SIGMA * (Th**4 - Tc**4) / (2.0 / eps - 1.0)
    assert q_i == 0.0
    assert abs(q_h - exact) <= 1e-9 * exact
    assert abs(q_c + exact) <= 1e-9 * exact


def test_isothermal_enclosure_has_zero_net_power():
    rng = np.random.default_rng(5)
    phases = (rng.random((10, 8, 8)) < 0.15).astype(np.uint8)
    P = O.build_surface_patches(phases, 0.01)
    o, r, c = O.first_hit_counts(P, 64, 21)
    op = X.OracleOperator(o, r, c, np.full(P.n_patches, 64), P.patch_area_m2)
    eps = np.full(P.n_patches, 0.7)
    E = X.emission(eps, P.patch_area_m2, np.full(P.n_patches, 600.0))
    absorbed, residual, _ = X.reflection_series(op, E, eps, relative_tolerance=1e-13)
    assert np.max(np.abs(absorbed - E)) <= 1e-9 * np.max(E)


def test_truncated_counts_raise_value_error():
    P = O.build_surface_patches(np.zeros((4, 3, 3), np.uint8), 0.004)
    o, r, c = O.first_hit_counts(P, 10, 1)
    with pytest.raises(ValueError, match="do not account for every emitted ray"):
        X.OracleOperator(o[:-1], r[:-1], c[:-1], np.full(P.n_patches, 10), P.patch_area_m2)
