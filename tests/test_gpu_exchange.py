"""Parity of the HIP exchange operator and FP64 reflection series with the CPU
oracle, the archived prescribed-field fluxes and closed forms.

Tolerances (FP64, different summation order from the reference):
  operator products        1e-13 relative to max|y|
  reflection series        1e-12 relative to max|absorbed| (same scale d)
  archived channels        1e-10 relative to max(|q_h|, |q_c|)
"""

from __future__ import annotations

import numpy as np
import pytest

from golden_data import cases, phases_x
from oracle import exchange_oracle as X
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mods():
    from porerad import _native, coupled, exchange, view_factors

    _native.lib()
    return view_factors, exchange, coupled


def _small(seed=2, shape=(14, 12, 12), density=0.15, R=64):
    rng = np.random.default_rng(seed)
    phases = (rng.random(shape) < density).astype(np.uint8)
    Q = O.build_surface_patches(phases, 0.01)
    o, r, c = O.first_hit_counts(Q, R, 31)
    return phases, Q, (o, r, c), np.full(Q.n_patches, R, np.int64)


@pytest.fixture(scope="module")
def small_pair(mods):
    vf, ex, _ = mods
    phases, Q, (o, r, c), rays = _small()
    op = ex.ExchangeOperator(o, r, c, rays, Q.patch_area_m2)
    ref = X.OracleOperator(o, r, c, rays, Q.patch_area_m2)
    return Q, op, ref


def test_operator_products_match_oracle(small_pair):
    Q, op, ref = small_pair
    op.device.set_scale(ref.scale)
    op.scale = ref.scale
    rng = np.random.default_rng(0)
    x = rng.random(Q.n_patches) + 0.1
    for name in ("raw_matvec", "raw_rmatvec", "symmetric_matvec", "matvec", "rmatvec"):
        y, y_ref = getattr(op, name)(x), getattr(ref, name)(x)
        assert np.max(np.abs(y - y_ref)) <= 1e-13 * np.max(np.abs(y_ref)), name


def test_balancing_closes_rows_like_oracle(mods):
    _, ex, _ = mods
    phases, Q, (o, r, c), rays = _small(7)
    op = ex.ExchangeOperator(o, r, c, rays, Q.patch_area_m2)
    ref = X.OracleOperator(o, r, c, rays, Q.patch_area_m2)
    assert op.row_sum_error <= 1e-10
    assert op.balance_method == ref.balance_method
    assert np.max(np.abs(op.scale / ref.scale - 1.0)) <= 1e-8
    assert op.reciprocity_error() <= 1e-12


@pytest.mark.parametrize("nonuniform", [False, True])
def test_reciprocity_error_on_device_matches_oracle(mods, nonuniform):
    # the device evaluates the reference's expression on the stored symmetric
    # matrix; with the oracle's factor d the same pairs give the same defect
    _, ex, _ = mods
    phases, Q, (o, r, c), rays = _small(11)
    area = Q.patch_area_m2.copy()
    if nonuniform:
        heavy = (o % 3) == 0
        c = np.where(heavy, 2 * c, c)
        rays = np.where(np.arange(Q.n_patches) % 3 == 0, 2 * rays, rays)
        area = area * (1.0 + 0.25 * (np.arange(Q.n_patches) % 4))
    op = ex.ExchangeOperator(o, r, c, rays, area)
    ref = X.OracleOperator(o, r, c, rays, area)
    op.device.set_scale(ref.scale)
    op.scale = ref.scale
    got, want = op.reciprocity_error(), ref.reciprocity_error()
    assert want <= 1e-12
    assert abs(got - want) <= 1e-3 * want + 1e-300, (got, want)


@pytest.mark.parametrize("eps", [0.3, 0.6, 0.9, "mixed"])
def test_reflection_series_matches_oracle(small_pair, mods, eps):
    _, ex, _ = mods
    Q, op, ref = small_pair
    op.device.set_scale(ref.scale)
    n = Q.n_patches
    rng = np.random.default_rng(1)
    e = rng.uniform(0.2, 0.95, n) if eps == "mixed" else np.full(n, eps)
    E = rng.normal(size=n) * 1e-3
    res = ex.reflection_series(op, E, e, relative_tolerance=1e-13)
    absorbed, residual, iters = X.reflection_series(ref, E, e, relative_tolerance=1e-13)
    assert abs(res.iterations - iters) <= 1
    assert np.max(np.abs(res.absorbed_power_W - absorbed)) <= 1e-12 * np.max(np.abs(absorbed))
    assert abs(res.global_energy_residual_W) <= 1e-12 * np.sum(np.abs(E))


def test_batched_series_equals_single_bitwise(small_pair, mods):
    _, ex, _ = mods
    Q, op, _ = small_pair
    rng = np.random.default_rng(3)
    E = rng.normal(size=(3, Q.n_patches))
    eps = np.full(Q.n_patches, 0.5)
    batch = ex.reflection_series_batch(op, E, eps, relative_tolerance=1e-8)
    for q in range(3):
        single = ex.reflection_series(op, E[q], eps, relative_tolerance=1e-8)
        assert np.array_equal(single.absorbed_power_W, batch[q].absorbed_power_W)
        assert single.iterations == batch[q].iterations


def _pf_params():
    return [pytest.param(c, id=c["case_id"]) for c in cases()["prescribed_field"]]


@pytest.mark.parametrize("case", _pf_params())
def test_prescribed_field_channels_match_archive(mods, case):
    vf, ex, cp = mods
    P = vf.build_surface_patches(phases_x(case["microstructure"], case["axis"]), case["H_m"])
    R = case["rays_per_patch_per_block"]
    o, r, c = vf.first_hit_counts_pooled(P, R, case["final_seeds"])
    rays = np.full(P.n_patches, R * len(case["final_seeds"]), np.int64)
    op = ex.ExchangeOperator(o, r, c, rays, P.patch_area_m2)
    assert op.row_sum_error <= 1e-10
    for ref in case["results"]:
        ch = cp.prescribed_field_channels(op, P, ref["T_mean_K"], ref["delta_T_K"], ref["epsilon"])
        scale = max(abs(ref["q_h_W_m2"]), abs(ref["q_c_W_m2"]))
        for got, want in ((ch.q_h, ref["q_h_W_m2"]), (ch.q_c, ref["q_c_W_m2"]), (ch.q_i, ref["q_i_W_m2"])):
            assert abs(got - want) <= 1e-10 * scale, (ref, got, want)
        assert ch.closure_rel <= 1e-4


@pytest.mark.parametrize("eps", [0.3, 0.9])
def test_empty_box_is_parallel_plates(mods, eps):
    vf, ex, cp = mods
    P = vf.build_surface_patches(np.zeros((6, 5, 5), np.uint8), 0.006)
    o, r, c = vf.first_hit_counts(P, 40, 11)
    op = ex.ExchangeOperator(o, r, c, np.full(P.n_patches, 40), P.patch_area_m2)
    ch = cp.prescribed_field_channels(op, P, 500.0, 400.0, eps)
    exact = X.SIGMA * (700.0**4 - 300.0**4) / (2.0 / eps - 1.0)
    assert abs(ch.q_h - exact) <= 1e-9 * exact and abs(ch.q_c + exact) <= 1e-9 * exact


def test_error_paths(mods, small_pair):
    _, ex, _ = mods
    phases, Q, (o, r, c), rays = _small(9)
    with pytest.raises(ValueError, match="do not account for every emitted ray"):
        ex.ExchangeOperator(o[:-1], r[:-1], c[:-1], rays, Q.patch_area_m2)
    _, op, _ = small_pair
    E = np.ones(op.n_patches)
    with pytest.raises(RuntimeError, match="did not converge"):
        ex.reflection_series(op, E, np.full(op.n_patches, 0.3), relative_tolerance=1e-13, max_reflections=1)


@pytest.mark.parametrize("uniform", [True, False])
def test_device_assembly_equals_host_assembly(mods, uniform):
    """prh_operator_create_device from the tracer's device edges gives the host
    build's operator bit for bit: count table, balancing, every product and
    the reflection series."""

    vf, ex, _ = mods
    rng = np.random.default_rng(7)
    phases = (rng.random((16, 13, 11)) < 0.18).astype(np.uint8)
    P = vf.build_surface_patches(phases, 0.01)
    R = 48
    area = P.patch_area_m2.copy()
    if not uniform:
        area *= 1.0 + 0.25 * rng.random(P.n_patches)  # areas that differ: the FP64 / FP32 forms
    rays = np.full(P.n_patches, 2 * R, np.int64)
    o, r, c = vf.first_hit_counts_pooled(P, R, [3, 4])
    host = ex.ExchangeOperator(o, r, c, rays, area)
    edges = vf.first_hit_edges(P, R, [3, 4])
    try:
        dev = ex.ExchangeOperator.from_edges(edges, rays, area)
    finally:
        edges.close()
    assert dev.device.nnz_symmetric == host.device.nnz_symmetric and dev.device.nnz_raw == host.device.nnz_raw
    for a, b in zip((dev.origin, dev.receiver, dev.count), (host.origin, host.receiver, host.count)):
        assert np.array_equal(a, b)
    assert np.array_equal(dev.weight, host.weight)
    assert dev.balance_iterations == host.balance_iterations and np.array_equal(dev.scale, host.scale)
    x = rng.random(P.n_patches)
    for name in ("raw_matvec", "raw_rmatvec", "symmetric_matvec", "matvec", "rmatvec"):
        assert np.array_equal(getattr(dev, name)(x), getattr(host, name)(x)), name
    E = rng.random(P.n_patches)
    e = np.full(P.n_patches, 0.6)
    a = ex.reflection_series(dev, E, e, relative_tolerance=1e-13)
    b = ex.reflection_series(host, E, e, relative_tolerance=1e-13)
    assert a.iterations == b.iterations and np.array_equal(a.absorbed_power_W, b.absorbed_power_W)


def test_device_assembly_checks_row_totals(mods):
    vf, ex, _ = mods
    phases = np.zeros((6, 5, 5), np.uint8)
    phases[2:4, 1:3, 1:3] = 1
    P = vf.build_surface_patches(phases, 0.005)
    edges = vf.first_hit_edges(P, 20, [1])
    try:
        with pytest.raises(ValueError, match="first-hit counts do not account for every emitted ray"):
            ex.ExchangeOperator.from_edges(edges, np.full(P.n_patches, 21, np.int64), P.patch_area_m2)
    finally:
        edges.close()
