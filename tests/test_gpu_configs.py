"""The benchmark workloads at full size, through the C ABI.

  random fibres N=64, 200 rays/patch (seed 1)      counts bit-exact against the
                                                   CPU oracle (count_sha256,
                                                   edges, occupancy reads);
                                                   FP64 channels at eps 0.6
  Voronoi foam N=128, 1000 rays/patch (seed 3),    the same along x, y and z;
  axes x, y, z                                     FP64 channels at eps 0.3,
                                                   0.6, 0.9
  foam + aligned fibres N=128                      served mixed-precision
                                                   channels (FP32 sweeps + FP64
                                                   refinement to 1e-10) within
                                                   1e-8 of the FP64 series,
                                                   per flux
  16 geometries N=96..192 (seeds 10..25)           one replayed batch pass ==
                                                   16 separate calls, edge for
                                                   edge (compared on the device)
  Poisson stream (200/s, seed 7) of radiosity and  every reply the same alone
  Newton-Krylov requests over the 16 geometries    and in the mix

The oracle's numbers (tests/golden/configs.json, from
tests/golden/make_config_fixtures.py) are porerad/view_factors.py:273-323
and porerad/exchange.py:47-245 + coupled.py:170-194 restated on the CPU.
FP64 channels are held to 1e-10 of max(|q_h|, |q_c|): the device balances
the operator with its own restatement of SciPy's Anderson mixing, which
agrees with SciPy's to round-off, not bit for bit.
"""

from __future__ import annotations

import numpy as np
import pytest

from golden_data import configs
from oracle import oracle as O

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10
MIXED_TOL = 1e-8


@pytest.fixture(scope="module")
def mods():
    from porerad import _native, coupled, exchange, serving, synthetic, view_factors

    _native.lib()
    return _native, view_factors, exchange, coupled, serving, synthetic


def _phase_sha(phases):
    import hashlib

    return hashlib.sha256(np.ascontiguousarray(phases, np.uint8).tobytes()).hexdigest()


_operators: dict = {}


def _traced(mods, name, phases, R, seed, fixture):
    """Trace one workload, check its counts against the oracle's, return (patches, operator)."""

    _native, vf, ex, _, _, _ = mods
    assert _phase_sha(phases) == fixture["phase_sha256"], "the generator no longer gives the oracle's geometry"
    P = vf.build_surface_patches(phases, configs()["H_m"])
    assert P.n_patches == fixture["n_patches"]
    edges = vf.first_hit_edges(P, R, [seed])
    try:
        assert edges.nnz == fixture["nnz"]
        assert edges.occupancy_reads == fixture["occupancy_reads"]
        o, r, c = edges.to_host()
        rays = np.full(P.n_patches, R, np.int64)
        assert vf.count_sha256(o, r, c, rays) == fixture["count_sha256"]
        del o, r, c
        op = ex.ExchangeOperator.from_edges(edges, rays, P.patch_area_m2)
    finally:
        edges.close()
    _operators[name] = (P, op)
    return P, op


def _check_channels(mods, P, op, fixture):
    _, _, _, coupled, _, _ = mods
    c = configs()
    T_m = 0.5 * (c["T_hot_K"] + c["T_cold_K"])
    for eps, ref in fixture["channels"].items():
        ch = coupled.prescribed_field_channels(op, P, T_m, c["T_hot_K"] - c["T_cold_K"], float(eps))
        scale = max(abs(ref["q_h"]), abs(ref["q_c"]))
        for q in ("q_h", "q_c", "q_i"):
            err = abs(getattr(ch, q) - ref[q]) / scale
            assert err <= FP64_TOL, (eps, q, err)


def test_fibre_n64_counts_and_channels(mods):
    synthetic = mods[5]
    fx = configs()["fibre_n64"]["x"]
    P, op = _traced(mods, "fibre_n64", synthetic.fibre_n64(), 200, 1, fx)
    _check_channels(mods, P, op, fx)
    _operators.pop("fibre_n64")[1].device.close()


@pytest.mark.parametrize("axis", ["x", "y", "z"])
def test_foam_n128_counts_and_channels(mods, axis):
    synthetic = mods[5]
    fx = configs()["foam_n128"][axis]
    P, op = _traced(mods, f"foam_n128_{axis}", O.orient_along_x(synthetic.foam_n128(), axis), 1000, 3, fx)
    _check_channels(mods, P, op, fx)
    if axis != "x":
        _operators.pop(f"foam_n128_{axis}")[1].device.close()


def test_mixed_precision_per_flux_at_full_size(mods):
    """Foam and aligned fibres N=128: served FP32/FP64-refined channels against the FP64 series."""

    _native, vf, ex, coupled, serving, synthetic = mods
    if "foam_n128_x" not in _operators:
        pytest.skip("needs the foam operator of test_foam_n128_counts_and_channels[x]")
    P, op = vf.build_surface_patches(synthetic.aligned_fibre_n128(), 0.01), None
    edges = vf.first_hit_edges(P, 1000, [3])
    try:
        op = ex.ExchangeOperator.from_edges(edges, np.full(P.n_patches, 1000, np.int64), P.patch_area_m2)
    finally:
        edges.close()
    _operators["aligned_n128"] = (P, op)
    srv = serving.RadiosityServer()
    for key in ("foam_n128_x", "aligned_n128"):
        srv.add_geometry(key, *_operators[key])
    combos = [(Th, eps) for Th in (500.0, 700.0, 900.0) for eps in (0.3, 0.6, 0.9)]
    reqs = [serving.RadiosityRequest(k, Th, 300.0, eps, 1e-10) for k in ("foam_n128_x", "aligned_n128")
            for Th, eps in combos]
    replies = srv.run(reqs, real_time=False)
    srv.close()
    assert len(replies) == len(reqs)
    worst = 0.0
    for rep in replies:
        Pk, opk = _operators[rep.request.key]
        Th = rep.request.T_hot_K
        ch = coupled.prescribed_field_channels(opk, Pk, 0.5 * (Th + 300.0), Th - 300.0, rep.request.epsilon)
        scale = max(abs(ch.q_h), abs(ch.q_c))
        for q in ("q_h", "q_c", "q_i"):
            err = abs(getattr(rep, q) - getattr(ch, q)) / scale
            worst = max(worst, err)
            assert err <= MIXED_TOL, (rep.request, q, err)
    print(f"largest relative flux error of the served mixed-precision solves: {worst:.2e}")
    for key in ("foam_n128_x", "aligned_n128"):
        _operators.pop(key)[1].device.close()


@pytest.fixture(scope="module")
def mixed_world(mods):
    """The 16 geometries traced in one replayed batch pass and geometry by geometry."""

    _native, vf, ex, _, serving, synthetic = mods
    geos = [(name, vf.build_surface_patches(phases, 0.01)) for name, phases in synthetic.mixed_geometries()]
    seeds = [[10 + k] for k in range(len(geos))]
    batch = _native.TraceBatch([(P.geometry(), 0, P.n_patches, 1000, 1) for _, P in geos])
    world = {"geos": geos, "seeds": seeds, "batch": batch, "operators": {}}
    yield world
    if world["batch"] is not None:
        world["batch"].close()
    for op in world["operators"].values():
        op.device.close()


def test_mixed_geometries_batch_equals_separate_calls(mods, mixed_world):
    _native, vf, ex, _, _, _ = mods
    geos, seeds, batch = mixed_world["geos"], mixed_world["seeds"], mixed_world["batch"]
    batch.run(seeds)
    tables = []
    try:
        for k, (name, P) in enumerate(geos):
            single = vf.first_hit_edges(P, 1000, seeds[k])
            tables.append(batch.edge_table(k))
            try:
                assert int(batch.nnz[k]) == single.nnz, name
                assert int(batch.occupancy_reads[k]) == single.occupancy_reads, name
                assert tables[-1].equals(single), name
            finally:
                single.close()
        # the batch's record buffers and the tracer's workspace go before the operators are built
        batch.close()
        mixed_world["batch"] = None
        _native.trim_tracer()
        for k, (name, P) in enumerate(geos):
            rays = np.full(P.n_patches, 1000, np.int64)
            mixed_world["operators"][name] = ex.ExchangeOperator.from_edges(tables[k], rays, P.patch_area_m2)
            tables[k].close()
    finally:
        for t in tables:
            t.close()


def test_poisson_stream_over_mixed_geometries(mods, mixed_world):
    """Radiosity and Newton-Krylov requests at 200/s over the 16 geometries: alone == in the mix."""

    _, _, _, _, serving, _ = mods
    ops = mixed_world["operators"]
    if len(ops) != 16:
        pytest.skip("needs the operators of test_mixed_geometries_batch_equals_separate_calls")
    patches = dict(mixed_world["geos"])

    def server():
        srv = serving.RadiosityServer()
        for name, op in ops.items():
            srv.add_geometry(name, patches[name], op)
        return srv

    rng = np.random.default_rng(7)
    names = list(ops)
    n_rad = 48
    arrivals = serving.poisson_arrivals(200.0, n_rad + 2, 7)
    reqs = [serving.RadiosityRequest(str(rng.choice(names)), float(rng.choice([500.0, 700.0, 900.0])), 300.0,
                                     float(rng.choice([0.3, 0.6, 0.9])), 1e-9, float(arrivals[i]), tag=i,
                                     patch_power=i < 8)
            for i in range(n_rad)]
    creqs = [serving.CoupledRequest("foam96", 700.0, 300.0, 0.6, 1.0, 0.026, float(arrivals[n_rad]), tag="c0",
                                    preconditioner_steps=(80,)),
             serving.CoupledRequest("random_fibre96", 500.0, 300.0, 0.3, 1.0, 0.026, float(arrivals[n_rad + 1]),
                                    tag="c1", preconditioner_steps=(80,))]
    srv = server()
    mixed = {r.request.tag: r for r in srv.run(reqs + creqs, real_time=True)}
    srv.close()
    assert len(mixed) == n_rad + 2
    for req in reqs[:8] + creqs:
        srv = server()
        alone = srv.run([req], real_time=False)[0]
        srv.close()
        got = mixed[req.tag]
        if isinstance(req, serving.CoupledRequest):
            assert alone.converged and got.converged and not got.failure
            assert got.newton_updates == alone.newton_updates
            assert got.metrics == alone.metrics
        else:
            assert np.array_equal(got.net_power_W, alone.net_power_W)
            assert (got.q_h, got.q_c, got.q_i, got.sweeps) == (alone.q_h, alone.q_c, alone.q_i, alone.sweeps)
