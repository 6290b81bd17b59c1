"""Served radiosity requests on the GPU: each result equals the FP64 reference
series (coupled.prescribed_field_channels, rtol 1e-13) within 1e-8 of
max(|q_h|, |q_c|), and is bit-identical whether it ran alone or in a mix of
requests on several resident geometries."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def world():
    from porerad import _native, coupled, exchange, serving, synthetic, view_factors

    _native.lib()
    geoms = {}
    for key, phases in (("fibre24", synthetic.random_fibers(24, 0.12, 2.0, 3)),
                        ("plates20", synthetic.layered_plates(20, 0.10, 11))):
        P = view_factors.build_surface_patches(phases, 0.005)
        o, r, c = view_factors.first_hit_counts_pooled(P, 150, [4, 5])
        op = exchange.ExchangeOperator(o, r, c, np.full(P.n_patches, 300, np.int64), P.patch_area_m2)
        geoms[key] = (P, op)
    return serving, coupled, geoms


def _server(serving, geoms):
    srv = serving.RadiosityServer()
    for key, (P, op) in geoms.items():
        srv.add_geometry(key, P, op)
    return srv


def test_served_channels_match_fp64_series(world):
    serving, coupled, geoms = world
    srv = _server(serving, geoms)
    reqs = [serving.RadiosityRequest(k, Th, 300.0, eps) for k in geoms for Th in (500.0, 900.0) for eps in (0.3, 0.9)]
    replies = srv.run(reqs, real_time=False)
    srv.close()
    assert len(replies) == len(reqs)
    for rep in replies:
        P, op = geoms[rep.request.key]
        T_m = 0.5 * (rep.request.T_hot_K + 300.0)
        ch = coupled.prescribed_field_channels(op, P, T_m, rep.request.T_hot_K - 300.0, rep.request.epsilon)
        scale = max(abs(ch.q_h), abs(ch.q_c))
        assert abs(rep.q_h - ch.q_h) <= 1e-8 * scale
        assert abs(rep.q_c - ch.q_c) <= 1e-8 * scale
        assert abs(rep.q_i - ch.q_i) <= 1e-8 * scale


def test_alone_equals_mixed_stream(world):
    serving, _, geoms = world
    rng = np.random.default_rng(7)
    specs = [(str(rng.choice(list(geoms))), float(rng.choice([500.0, 700.0, 900.0])), float(rng.choice([0.3, 0.6, 0.9])))
             for _ in range(40)]
    alone = {}
    for i, (k, Th, eps) in enumerate(specs[:6]):
        srv = _server(serving, geoms)
        alone[i] = srv.run([serving.RadiosityRequest(k, Th, 300.0, eps, tag=i, patch_power=True)],
                           real_time=False)[0]
        srv.close()
    srv = _server(serving, geoms)
    arrivals = serving.poisson_arrivals(2000.0, len(specs), 7)
    reqs = [serving.RadiosityRequest(k, Th, 300.0, eps, arrival_s=float(t), tag=i, patch_power=i < 6)
            for i, ((k, Th, eps), t) in enumerate(zip(specs, arrivals))]
    replies = srv.run(reqs, real_time=True)
    srv.close()
    assert len(replies) == len(specs)
    by_tag = {r.request.tag: r for r in replies}
    for i, rep in alone.items():
        assert np.array_equal(by_tag[i].net_power_W, rep.net_power_W)
        assert (by_tag[i].q_h, by_tag[i].q_c, by_tag[i].q_i) == (rep.q_h, rep.q_c, rep.q_i)
        assert by_tag[i].sweeps == rep.sweeps
    assert all(r.net_power_W is None for r in replies if r.request.tag >= 6)
    lat = serving.latency_summary(replies, 1.0)
    assert lat["p99_ms"] >= lat["p50_ms"] > 0


def test_coupled_requests_served_beside_radiosity(world):
    """Coupled solves on the worker thread give the bits of a direct solve, while radiosity requests keep flowing."""

    serving, coupled, geoms = world
    from porerad import conduction

    creqs = [serving.CoupledRequest("fibre24", 700.0, 300.0, 0.6, conductivity_solid=1.0, arrival_s=0.002, tag="c0"),
             serving.CoupledRequest("plates20", 900.0, 300.0, 0.3, conductivity_solid=1.0, arrival_s=0.004, tag="c1")]
    direct = {}
    for r in creqs:
        P, op = geoms[r.key]
        system = conduction.assemble_conduction(np.where(P.solid == 1, r.conductivity_solid, r.conductivity_gas),
                                                P.H_m)
        T0 = conduction.solve_conduction(system, r.T_hot_K, r.T_cold_K)
        res = coupled.solve_coupled(system, op, P, r.T_hot_K, r.T_cold_K, r.epsilon, T0)
        assert res.converged
        direct[r.tag] = coupled.wall_flux_metrics(system, P, res, r.T_hot_K, r.T_cold_K)
    rng = np.random.default_rng(3)
    arrivals = serving.poisson_arrivals(2000.0, 30, 3)
    reqs = [serving.RadiosityRequest(str(rng.choice(list(geoms))), 700.0, 300.0, float(rng.choice([0.3, 0.9])),
                                     arrival_s=float(t), tag=i) for i, t in enumerate(arrivals)]
    srv = _server(serving, geoms)
    replies = srv.run(reqs + creqs, real_time=True)
    srv.close()
    assert len(replies) == len(reqs) + len(creqs)
    served = {r.request.tag: r for r in replies if isinstance(r, serving.CoupledReply)}
    assert set(served) == {"c0", "c1"}
    for tag, m in direct.items():
        got = served[tag]
        assert got.converged and not got.failure
        for key in ("q_cond_hot_W_m2", "q_rad_hot_W_m2", "q_total_cold_W_m2", "T_min_K", "T_max_K"):
            assert got.metrics[key] == m[key], key
