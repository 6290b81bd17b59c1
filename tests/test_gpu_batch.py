"""The batched mixed-precision radiosity engine (prh_batch_*) against the FP64 oracle.

Tolerances: absorbed power within 1e-8 of max|absorbed| and channels within
1e-8 of max(|q_h|, |q_c|) of the converged FP64 series (oracle, rtol 1e-13),
at solver tolerance 1e-9.  A solve must return identical bits alone, in a
mix of other solves, and when admitted while others are running.
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
    return _native, view_factors, exchange, coupled


def _pair(seed=4, shape=(16, 14, 14), density=0.14, R=64, nonuniform=False):
    from porerad import exchange

    rng = np.random.default_rng(seed)
    phases = (rng.random(shape) < density).astype(np.uint8)
    Q = O.build_surface_patches(phases, 0.01)
    o, r, c = O.first_hit_counts(Q, R, 91)
    rays = np.full(Q.n_patches, R, np.int64)
    if nonuniform:
        # doubling every count of some origins keeps each row a valid ray set
        heavy = (o % 3) == 0
        c = np.where(heavy, 2 * c, c)
        rays = np.where(np.arange(Q.n_patches) % 3 == 0, 2 * R, R)
    op = exchange.ExchangeOperator(o, r, c, rays, Q.patch_area_m2)
    ref = X.OracleOperator(o, r, c, rays, Q.patch_area_m2)
    op.device.set_scale(ref.scale)
    op.scale = ref.scale
    return Q, op, ref


@pytest.mark.parametrize("nonuniform", [False, True])
@pytest.mark.parametrize("eps", [0.3, 0.6, 0.9, "mixed"])
def test_mixed_matches_fp64_oracle(mods, eps, nonuniform):
    _, _, ex, _ = mods
    Q, op, ref = _pair(nonuniform=nonuniform)
    assert op.device.nnz_symmetric > 0
    n = Q.n_patches
    rng = np.random.default_rng(5)
    e = rng.uniform(0.2, 0.95, n) if eps == "mixed" else np.full(n, eps)
    E = rng.normal(size=n) * 1e-3 + 2e-3
    got = ex.reflection_series_batch(op, E[None, :], e, relative_tolerance=1e-9, precision="mixed")[0]
    absorbed, _, _ = X.reflection_series(ref, E, e, relative_tolerance=1e-13)
    assert np.max(np.abs(got.absorbed_power_W - absorbed)) <= 1e-8 * np.max(np.abs(absorbed))
    assert abs(got.global_energy_residual_W) <= 1e-8 * np.sum(np.abs(E))


def test_alone_equals_in_mix_bitwise(mods):
    _native, _, ex, _ = mods
    Q, op, _ = _pair(6)
    n = Q.n_patches
    rng = np.random.default_rng(8)
    E = rng.normal(size=(5, n))
    eps = np.stack([np.full(n, v) for v in (0.3, 0.9, 0.6, 0.45, 0.75)])
    alone = [ex.reflection_series_batch(op, E[q], eps[q], relative_tolerance=1e-9, precision="mixed")[0]
             for q in range(5)]
    mixed = ex.reflection_series_batch(op, E, eps, relative_tolerance=1e-9, precision="mixed")
    for a, b in zip(alone, mixed):
        assert np.array_equal(a.absorbed_power_W, b.absorbed_power_W)
        assert a.iterations == b.iterations
    # admitted while others run
    batch = _native.Batch(op.device)
    try:
        s0 = batch.admit(E[0], eps[0], 1e-9)
        s1 = batch.admit(E[1], eps[1], 1e-9)
        batch.step(5)
        s2 = batch.admit(E[2], eps[2], 1e-9)
        results = {}
        while len(results) < 3:
            done, active = batch.step(64)
            assert active
            for s in (s0, s1, s2):
                if (done >> s) & 1 and s not in results:
                    results[s] = batch.result(s)
        for q, s in enumerate((s0, s1, s2)):
            assert np.array_equal(results[s][0], alone[q].absorbed_power_W)
    finally:
        batch.close()


@pytest.mark.parametrize("nonuniform", [False, True])
def test_concentrated_and_wide_range_emissions(mods, nonuniform):
    """Emission on one patch, and emissions spanning 30 decades across patches:
    the FP16 gathered directions stay in range (the first epoch is scaled by
    max|E|) and the solves meet the same 1e-8 bar against the oracle."""

    _native, _, ex, _ = mods
    Q, op, ref = _pair(12, nonuniform=nonuniform)
    n = Q.n_patches
    rng = np.random.default_rng(13)
    single = np.zeros(n)
    single[n // 3] = 5.0
    wide = rng.normal(size=n) * 10.0 ** rng.uniform(-20, 10, n)
    E = np.stack([single, wide])
    eps = np.stack([np.full(n, 0.4), rng.uniform(0.2, 0.95, n)])
    got = ex.reflection_series_batch(op, E, eps, relative_tolerance=1e-9, precision="mixed")
    for q in range(2):
        absorbed, _, _ = X.reflection_series(ref, E[q], eps[q], relative_tolerance=1e-13)
        assert np.max(np.abs(got[q].absorbed_power_W - absorbed)) <= 1e-8 * np.max(np.abs(absorbed)), q
        assert abs(got[q].global_energy_residual_W) <= 1e-8 * np.sum(np.abs(E[q])), q


@pytest.mark.parametrize("which", ["morton", "random"])
def test_sweep_order_keeps_results(mods, which):
    """Rows swept in another order: the same 1e-8 bar against the oracle, and a
    solve still returns the same bits alone and in a mix."""

    _native, vf, ex, _ = mods
    Q, op, ref = _pair(15)
    n = Q.n_patches
    if which == "morton":
        P = vf.build_surface_patches((np.random.default_rng(15).random((16, 14, 14)) < 0.14).astype(np.uint8), 0.01)
        assert P.n_patches == n
        order = P.sweep_order()
        assert np.array_equal(np.sort(order), np.arange(n))
    else:
        order = np.random.default_rng(16).permutation(n).astype(np.int32)
    op.set_sweep_order(order)
    rng = np.random.default_rng(17)
    E = rng.normal(size=(4, n)) * 1e-3
    eps = rng.uniform(0.25, 0.95, size=(4, n))
    mixed = ex.reflection_series_batch(op, E, eps, relative_tolerance=1e-9, precision="mixed")
    for q in range(4):
        absorbed, _, _ = X.reflection_series(ref, E[q], eps[q], relative_tolerance=1e-13)
        assert np.max(np.abs(mixed[q].absorbed_power_W - absorbed)) <= 1e-8 * np.max(np.abs(absorbed)), q
    alone = ex.reflection_series_batch(op, E[2], eps[2], relative_tolerance=1e-9, precision="mixed")[0]
    assert np.array_equal(alone.absorbed_power_W, mixed[2].absorbed_power_W)
    with pytest.raises(ValueError):
        op.set_sweep_order(np.zeros(n, np.int32))
    op.set_sweep_order(None)


def test_engine_bytes_match_serving_estimate(mods):
    """The server's HBM admission counts exactly what an engine allocates."""

    _native, vf, _, _ = mods
    from porerad import serving

    Q, op, _ = _pair(14, shape=(12, 10, 10))
    batch = _native.Batch(op.device)
    try:
        P = vf.build_surface_patches((np.random.default_rng(14).random((12, 10, 10)) < 0.14).astype(np.uint8), 0.01)
        g = serving.Geometry("g", P, op)
        assert P.n_patches == Q.n_patches
        batch.set_field(P.patch_class, g.cell_x, 12, g.cross_section)
        assert batch.bytes == serving.device_engine_bytes(g)
    finally:
        batch.close()


def test_more_requests_than_slots(mods):
    _native, _, ex, _ = mods
    Q, op, ref = _pair(7, shape=(12, 10, 10))
    n = Q.n_patches
    rng = np.random.default_rng(9)
    k = _native.BATCH_SLOTS + 7
    E = rng.normal(size=(k, n))
    eps = rng.uniform(0.25, 0.95, size=(k, n))
    got = ex.reflection_series_batch(op, E, eps, relative_tolerance=1e-9, precision="mixed")
    for q in (0, k // 2, k - 1):
        absorbed, _, _ = X.reflection_series(ref, E[q], eps[q], relative_tolerance=1e-13)
        assert np.max(np.abs(got[q].absorbed_power_W - absorbed)) <= 1e-8 * np.max(np.abs(absorbed))


def test_channels_of_archived_case_within_1e8(mods):
    _, vf, ex, cp = mods
    case = next(c for c in cases()["prescribed_field"] if c["case_id"] == "pf_OF_core_phi0.05_s3_x")
    P = vf.build_surface_patches(phases_x(case["microstructure"], case["axis"]), case["H_m"])
    R = case["rays_per_patch_per_block"]
    o, r, c = vf.first_hit_counts_pooled(P, R, case["final_seeds"])
    op = ex.ExchangeOperator(o, r, c, np.full(P.n_patches, R * 2, np.int64), P.patch_area_m2)
    assert op.device.nnz_symmetric > 0
    rows = case["results"]
    T_field = []
    E, eps = [], []
    for ref in rows:
        T = cp.linear_cell_temperature(P.solid.shape, ref["T_mean_K"], ref["delta_T_K"])
        T_patch = cp.patch_temperature(P, T, ref["T_mean_K"] + ref["delta_T_K"] / 2, ref["T_mean_K"] - ref["delta_T_K"] / 2)
        e = np.full(P.n_patches, ref["epsilon"])
        E.append(ex.emission(e, P.patch_area_m2, T_patch, ref["T_mean_K"]))
        eps.append(e)
        T_field.append(T)
    got = ex.reflection_series_batch(op, np.stack(E), np.stack(eps), relative_tolerance=1e-9, precision="mixed")
    for ref, res in zip(rows, got):
        q_h, q_c, q_i, _ = cp._channels_from_net(P, res.net_power_W)
        scale = max(abs(ref["q_h_W_m2"]), abs(ref["q_c_W_m2"]))
        for a, b in ((q_h, ref["q_h_W_m2"]), (q_c, ref["q_c_W_m2"]), (q_i, ref["q_i_W_m2"])):
            assert abs(a - b) <= 1e-8 * scale


def test_linear_field_admission_matches_host_emission(mods):
    """Channels of a device-formed linear field agree with the same solve of the host-formed emission."""

    _native, vf, ex, cp = mods
    from porerad import serving

    rng = np.random.default_rng(12)
    phases = (rng.random((18, 12, 12)) < 0.15).astype(np.uint8)
    P = vf.build_surface_patches(phases, 0.004)
    o, r, c = vf.first_hit_counts_pooled(P, 96, [21])
    op = ex.ExchangeOperator(o, r, c, np.full(P.n_patches, 96, np.int64), P.patch_area_m2)
    g = serving.Geometry("g", P, op)
    batch = _native.Batch(op.device)
    try:
        batch.set_field(P.patch_class, g.cell_x, g.nx, g.cross_section)
        for Th, eps in ((900.0, 0.3), (500.0, 0.9), (300.0, 0.6)):
            req = serving.RadiosityRequest("g", Th, 300.0, eps)
            a = batch.admit_linear(Th, 300.0, eps, 1e-10)
            assert batch.run_until_done() >> a & 1
            q_dev = batch.channels(a)
            absorbed, _, _ = batch.result(a)
            E = g.emission(req)
            b = batch.admit(E, eps, 1e-10)
            assert batch.run_until_done() >> b & 1
            absorbed_host, _, _ = batch.result(b)
            scale = np.max(np.abs(E)) if Th != 300.0 else 1.0
            assert np.max(np.abs(absorbed - absorbed_host)) <= 1e-9 * scale
            q_host = g.channels(absorbed - E)
            for x, y in zip(q_dev[:3], q_host):
                assert abs(x - y) <= 1e-9 * max(abs(q_host[0]), abs(q_host[1]), 1e-300)
            if Th == 300.0:
                assert q_dev[:3] == (0.0, 0.0, 0.0) and not np.any(absorbed)
    finally:
        batch.close()


def test_admission_errors(mods):
    _native, _, _, _ = mods
    Q, op, _ = _pair(10, shape=(8, 6, 6))
    batch = _native.Batch(op.device)
    try:
        with pytest.raises(ValueError):
            batch.admit(np.ones(Q.n_patches), 0.0, 1e-9)  # nothing absorbs
        for _ in range(_native.BATCH_SLOTS):
            batch.admit(np.ones(Q.n_patches), 0.5, 1e-9)
        with pytest.raises(ValueError, match="slots are busy"):
            batch.admit(np.ones(Q.n_patches), 0.5, 1e-9)
        with pytest.raises(ValueError, match="no field"):
            batch.admit_linear(700.0, 300.0, 0.5, 1e-9)
        with pytest.raises(ValueError, match="no finished solve"):
            batch.release(0)
    finally:
        batch.close()


@pytest.mark.parametrize("nonuniform", [False, True])
@pytest.mark.parametrize("n_rhs", [5, 20, 40])
def test_fp64_batched_series_matches_oracle(mods, n_rhs, nonuniform):
    """The FP64 series of 5 (16-slot layout), 20 (32-slot) and 40 (32 + 8)
    emissions in batched passes: per emission within 1e-12 of the oracle's
    series, the same reflection count as the one-at-a-time device series
    (within one, the stopping sums differ in order), energy residual small."""

    _, _, ex, _ = mods
    Q, op, ref = _pair(nonuniform=nonuniform)
    n = Q.n_patches
    rng = np.random.default_rng(40 + n_rhs)
    E = rng.normal(size=(n_rhs, n)) * 1e-3
    E[1] = 0.0  # the zero emission stops after its first reflection
    eps = rng.uniform(0.2, 0.95, (n_rhs, n))
    got = ex.reflection_series_batch(op, E, eps, relative_tolerance=1e-12, precision="fp64_batched")
    single = ex.reflection_series_batch(op, E, eps, relative_tolerance=1e-12, precision="fp64")
    for q in range(n_rhs):
        absorbed, _, iters = X.reflection_series(ref, E[q], eps[q], relative_tolerance=1e-12)
        scale = max(np.max(np.abs(absorbed)), 1e-300)
        assert np.max(np.abs(got[q].absorbed_power_W - absorbed)) <= 1e-12 * scale, q
        assert abs(got[q].iterations - single[q].iterations) <= 1, q
        assert abs(got[q].iterations - iters) <= 1, q
        assert abs(got[q].global_energy_residual_W) <= 1e-11 * max(np.sum(np.abs(E[q])), 1e-300), q
    assert got[1].iterations == 1 and not np.any(got[1].absorbed_power_W)
