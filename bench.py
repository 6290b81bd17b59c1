"""Benchmark of the porerad radiation hot path on one MI355X.

Workload (one GPU): open-cell foam N=128^3 from the periodic Voronoi edges of
64 uniform seeds, strut radius bisected to solid fraction 0.10 (seed 2), axis
x, 1000 cosine-weighted rays per patch (seed 3).

A step is one fused first-hit pass over every patch of the foam: seeded ray
generation -> FP64 DDA -> per-patch integer counts -> sorted edge table in
HBM (prh_trace_counts).  ``value`` is first-hit rays per second of the whole
step with the geometry already resident; the step's count_sha256 is checked
against the CPU oracle's (``parity``).  Alongside it the line reports:

  radiosity    plain FP64 series and served mixed-precision solves on the
               foam and the aligned fibres N=128 (Poisson 200/s), saturation,
               error per flux, tick/series rooflines, coupled requests, and
               the C oracle's FP64 series on the foam operator (solves/s)
  batching     16 uneven geometries N=96..192 in one replayed graph pass
               (rays/s) against 16 separate calls, then a Poisson stream
               (200/s, seed 7) of radiosity and Newton-Krylov requests over
               their 16 resident operators (solves/s, p50/p99, coupled
               latencies)
  monte_carlo, microstructure   the rows ranked next to the path
  cpu_baseline the C oracle's first_hit_counts on this host's threads and
               on one core (CPU model and counts stated)

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""

from __future__ import annotations

import argparse
import dataclasses
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "pore-radiation-ijhmt2026_amd"))

import numpy as np  # noqa: E402

HBM_PEAK_GBPS = 8000.0
METRIC = "first-hit rays/s at N=128; served radiosity solves/s, p50/p99 at 1e-8 rel err"


def traffic_profile(kernel: str) -> str:
    return os.path.join(REPO, "profiles", f"{kernel}_foam128_hbm_traffic.md")


def traversal_kernel() -> str:
    """Name of the traversal kernel the current tracer variant launches on the foam."""

    from porerad import _native

    current = next(v for v in _native.tracer_variants() if v["current"])
    return "k_trace_lean" if current["occupancy"] == 3 else "k_trace_stream"


def counters_profile(kernel: str) -> str:
    return os.path.join(REPO, "profiles", f"{kernel}_foam128_counters.md")


def profile_value(path: str, key: str) -> float | None:
    """The value of a `key: value` line of a committed counter summary (None if absent)."""

    try:
        with open(path) as f:
            for line in f:
                if line.startswith(key + ":"):
                    return float(line.split(":")[1])
    except OSError:
        pass
    return None


def measured_traffic(kernel: str) -> float | None:
    """HBM bytes per traversal launch at the default workload, from the committed counter summary."""

    return profile_value(traffic_profile(kernel), "traffic_bytes_per_launch")


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def foam_workload():
    from porerad import synthetic

    return synthetic.foam_n128()


def run_tracer(P, R, seed, steps, warmup):
    from porerad import _native, view_factors

    for _ in range(warmup):
        view_factors.first_hit_edges(P, R, [seed]).close()
    wall, parts = [], []
    reads = nnz = 0
    for _ in range(steps):
        t0 = time.perf_counter()
        e = view_factors.first_hit_edges(P, R, [seed])
        wall.append(time.perf_counter() - t0)
        parts.append(_native.last_timings() * 1e-3)
        reads, nnz = e.occupancy_reads, e.nnz
        e.close()
    return np.array(wall), np.array(parts), reads, nnz


def lane_share(P, R, seed, reads):
    """Share of wavefront lanes that step in one traced pass of the lean traversal
    (occupancy reads over 64 x wave-steps, from the kernel's own counters)."""

    from porerad import _native, view_factors

    _native.tracer_stats(reset=True)
    view_factors.first_hit_edges(P, R, [seed]).close()
    st = _native.tracer_stats(reset=True)
    if st["wave_steps"] <= 0:
        return None
    return {"wave_steps": st["wave_steps"], "refills": st["refills"],
            "share": float(reads) / (64.0 * st["wave_steps"])}


def count_parity(P, R, seed):
    """count_sha256 of one traced step against the CPU oracle's for the same workload
    (tests/golden/configs.json, written by the oracle on the CPU); raises on a mismatch."""

    from porerad import view_factors

    path = os.path.join(REPO, "tests", "golden", "configs.json")
    with open(path) as f:
        fx = json.load(f)["foam_n128"]["x"]
    if (R, seed, P.n_patches) != (fx["rays_per_patch"], fx["seed"], fx["n_patches"]):
        return {"checked": False, "reason": "no oracle fixture for this ray count"}
    e = view_factors.first_hit_edges(P, R, [seed])
    try:
        o, r, c = e.to_host()
        reads = e.occupancy_reads
    finally:
        e.close()
    digest = view_factors.count_sha256(o, r, c, np.full(P.n_patches, R, np.int64))
    ok = digest == fx["count_sha256"] and o.size == fx["nnz"] and reads == fx["occupancy_reads"]
    if not ok:
        raise RuntimeError(f"first-hit counts differ from the CPU oracle's: {digest} vs {fx['count_sha256']}")
    return {"checked": True, "count_sha256": digest, "edges": int(o.size), "occupancy_reads": int(reads),
            "equal_to_cpu_oracle": True}


def build_operator(P, R, seed):
    """Counts traced and the operator assembled on the device (no edge leaves HBM)."""

    from porerad import exchange, view_factors

    edges = view_factors.first_hit_edges(P, R, [seed])
    try:
        op = exchange.ExchangeOperator.from_edges(edges, np.full(P.n_patches, R, np.int64), P.patch_area_m2)
    finally:
        edges.close()
    op.set_sweep_order(P.sweep_order())  # the batched engine sweeps rows along a Morton curve
    return op


def run_radiosity(foam_P, R, seed, n_requests=600, rate=200.0, with_cpu=True):
    """Plain FP64 series vs served mixed-precision solves (Poisson arrivals)."""

    from porerad import coupled, exchange, serving, synthetic, view_factors
    from porerad.constants import K_GAS

    t0 = time.perf_counter()
    geoms = {"foam128": (foam_P, build_operator(foam_P, R, seed))}
    fibre = view_factors.build_surface_patches(synthetic.aligned_fibre_n128(), 0.01)
    geoms["aligned128"] = (fibre, build_operator(fibre, R, seed))
    setup = time.perf_counter() - t0
    log(f"operators: {', '.join(f'{k}: {P.n_patches} patches, {op.device.nnz_symmetric} nnz' for k, (P, op) in geoms.items())}"
        f" ({setup:.1f} s)")
    combos = [(Th, eps) for Th in (500.0, 700.0, 900.0) for eps in (0.3, 0.6, 0.9)]

    # plain FP64 path: the reference algorithm (reflection series) per request
    P, op = geoms["foam128"]
    plain_lat, refs = [], {}
    series0 = op.device.stats()
    for Th, eps in combos:
        T_m = 0.5 * (Th + 300.0)
        t = time.perf_counter()
        st = coupled.evaluate_radiation(op, P, coupled.linear_cell_temperature(P.solid.shape, T_m, Th - 300.0),
                                        Th, 300.0, eps, relative_tolerance=1e-8, reference_temperature_K=T_m)
        plain_lat.append(time.perf_counter() - t)
        _ = st
    series1 = op.device.stats()
    plain = {"solves_per_s": len(plain_lat) / sum(plain_lat), "p50_ms": float(np.percentile(plain_lat, 50) * 1e3),
             "p99_ms": float(np.percentile(plain_lat, 99) * 1e3), "precision": "fp64 reflection series, rtol 1e-8",
             "geometry": "foam128"}
    cpu_rad = cpu_radiosity_baseline(P, op) if with_cpu else None
    if cpu_rad is not None:
        log(f"radiosity cpu baseline {cpu_rad['value']:.3f} solves/s ({cpu_rad['reflections']} reflections, "
            f"{cpu_rad['cores']} threads)")

    batching_vs_precision = precision_decomposition(geoms["foam128"])

    # FP64 reflection series at 1e-13 for the accuracy check, both geometries
    for key, (Pk, opk) in geoms.items():
        for Th, eps in combos:
            T_m = 0.5 * (Th + 300.0)
            refs[(key, Th, eps)] = coupled.prescribed_field_channels(opk, Pk, T_m, Th - 300.0, eps)

    def make_requests(count, seed_, rtol=1e-9):
        rng = np.random.default_rng(seed_)
        keys = list(geoms)
        arrivals = serving.poisson_arrivals(rate, count, seed_)
        return [serving.RadiosityRequest(str(rng.choice(keys)), float(rng.choice([500.0, 700.0, 900.0])), 300.0,
                                         float(rng.choice([0.3, 0.6, 0.9])), rtol, float(t)) for t in arrivals]

    def flux_errors(replies):
        err = {}
        for rep in replies:
            if not isinstance(rep, serving.RadiosityReply):
                continue
            ref = refs[(rep.request.key, rep.request.T_hot_K, rep.request.epsilon)]
            scale = max(abs(ref.q_h), abs(ref.q_c))
            e = err.setdefault(rep.request.key, {"q_h": 0.0, "q_c": 0.0, "q_i": 0.0})
            for name, got, want in (("q_h", rep.q_h, ref.q_h), ("q_c", rep.q_c, ref.q_c), ("q_i", rep.q_i, ref.q_i)):
                e[name] = max(e[name], abs(got - want) / scale)
        return err

    srv = serving.RadiosityServer()
    for k, (Pk, opk) in geoms.items():
        srv.add_geometry(k, Pk, opk)
    srv.run(make_requests(32, 99), real_time=False)  # warm-up: engines allocated
    reqs = make_requests(n_requests, 7)
    t = time.perf_counter()
    replies = srv.run(reqs, real_time=True)
    served = serving.latency_summary(replies, time.perf_counter() - t)
    served.update({"arrival_rate_per_s": rate, "requests": n_requests, "geometries": list(geoms),
                   "precision": "chebyshev sweeps (fp16 gathered directions, fp32 recursions) + fp64 refinement, rtol 1e-9", "sweeps": srv.sweeps})
    ticks0 = srv.engine_stats()
    t = time.perf_counter()
    sat = srv.run(make_requests(n_requests, 8), real_time=False)
    saturation = serving.latency_summary(sat, time.perf_counter() - t)
    ticks1 = srv.engine_stats()

    # precision: every (geometry, T_hot, eps) refined to 1e-10 against the FP64 series at 1e-13
    prec = [serving.RadiosityRequest(k, Th, 300.0, eps, 1e-10) for k in geoms for Th, eps in combos]
    precision = {"relative_tolerance": 1e-10, "reference": "fp64 reflection series, rtol 1e-13",
                 "max_rel_error_per_flux": flux_errors(srv.run(prec, real_time=False))}

    # serving plus coupling: radiosity requests at the same rate with coupled solves of the foam
    # arriving among them (solid 1.0 and gas 0.026 W/m/K; mixed-precision Jacobian solves)
    mixed_reqs = make_requests(n_requests // 2, 9)
    cases = ((500.0, 0.3), (700.0, 0.6), (900.0, 0.9))
    coupled_reqs = [serving.CoupledRequest("foam128", Th, 300.0, eps, C5_K_SOLID, K_GAS, arrival_s=a,
                                           precision="mixed", preconditioner_steps=C5_PCG_STEPS)
                    for (Th, eps), a in zip(cases, (0.05, 0.3, 0.6))]
    t = time.perf_counter()
    mixed = srv.run(mixed_reqs + coupled_reqs, real_time=True)
    rad_part = [r for r in mixed if isinstance(r, serving.RadiosityReply)]
    cpl_part = [r for r in mixed if isinstance(r, serving.CoupledReply)]
    rad_wall = max(r.finished_s for r in rad_part) - t
    coupled_block = {
        "radiosity": serving.latency_summary(rad_part, rad_wall),
        "coupled": {"solves": len(cpl_part), "latency_s": [round(r.latency_s, 3) for r in cpl_part],
                    "newton_updates": [r.newton_updates for r in cpl_part],
                    "converged": all(r.converged for r in cpl_part),
                    "q_total_hot_W_m2": [r.metrics["q_total_hot_W_m2"] for r in cpl_part],
                    "wall_flux_mismatch_rel": max(r.metrics["wall_flux_mismatch_rel"] for r in cpl_part)},
        "workload": f"{n_requests // 2} radiosity requests at {rate:.0f}/s + {len(cases)} coupled solves of the "
                    f"foam (k_solid {C5_K_SOLID}, k_gas {K_GAS} W/m/K, Jacobi-PCG {C5_PCG_STEPS[0]} steps, "
                    f"mixed-precision Jacobian solves)",
    }
    # each coupled request alone, with FP64 and with mixed-precision Jacobian solves: latency and
    # the error of the mixed path per wall flux
    alone = {}
    for prec in ("fp64", "mixed"):
        alone[prec] = {}
        for r in coupled_reqs:
            rep, = srv.run([dataclasses.replace(r, arrival_s=0.0, precision=prec)], real_time=False)
            alone[prec][(r.T_hot_K, r.epsilon)] = rep
    flux_keys = ("q_cond_hot_W_m2", "q_rad_hot_W_m2", "q_cond_cold_W_m2", "q_rad_cold_W_m2")
    cpl_err = {}
    for key, ref in alone["fp64"].items():
        got = alone["mixed"][key]
        scale = max(abs(ref.metrics["q_total_hot_W_m2"]), abs(ref.metrics["q_total_cold_W_m2"]))
        cpl_err[f"T_hot {key[0]:.0f} K, eps {key[1]}"] = {k: abs(got.metrics[k] - ref.metrics[k]) / scale
                                                          for k in flux_keys}
    coupled_block["coupled"]["alone_latency_s"] = {
        prec: [round(rep.latency_s, 3) for rep in alone[prec].values()] for prec in alone}
    coupled_block["coupled"]["alone_converged"] = all(rep.converged for d in alone.values() for rep in d.values())
    coupled_block["coupled"]["max_rel_flux_error_vs_fp64"] = cpl_err
    srv.close()
    err = flux_errors(replies + sat)
    roofline = {"engine": engine_roofline(ticks0, ticks1), "fp64_series": series_roofline(series0, series1)}
    return {
        "roofline": roofline,
        "plain_fp64": plain,
        "served": served,
        "saturation": saturation,
        "speedup_vs_plain_fp64": saturation["solves_per_s"] / plain["solves_per_s"],
        "batching_vs_precision": batching_vs_precision,
        "max_rel_flux_error_vs_fp64": {k: max(v.values()) for k, v in err.items()},
        "precision": precision,
        "serving_with_coupled": coupled_block,
        "operator_setup_s": round(setup, 2),
        "cpu_baseline": cpu_rad,
    }


def precision_decomposition(geometry, n_emissions=64, rtol=1e-9):
    """The same emissions solved four ways, to separate batching from precision.

    64 prescribed-field emissions of one geometry (T_hot 500..900 K, eps
    0.3..0.9) to ||flight||_1 <= 1e-9 ||E||_1: the FP64 series one emission
    at a time, the FP64 series 16 and 32 emissions per pass over the operator
    (FP64 slots interleaved per patch), and the mixed-precision engine (64
    slots: FP16 gathered directions, FP64 refinement).  Solves/s of each, the
    reflections or sweeps they made, and the largest flux error against the
    one-at-a-time FP64 series."""

    from porerad import exchange, serving

    P, op = geometry
    g = serving.Geometry("decomposition", P, op)
    rng = np.random.default_rng(12)
    reqs = [serving.RadiosityRequest("decomposition", float(rng.uniform(500.0, 900.0)), 300.0,
                                     float(rng.choice([0.3, 0.6, 0.9])), rtol) for _ in range(n_emissions)]
    E = np.stack([g.emission(r) for r in reqs])
    eps = np.stack([np.full(g.n, r.epsilon) for r in reqs])

    def solve(precision, group):
        out = []
        for q0 in range(0, n_emissions, group):
            out += exchange.reflection_series_batch(op, E[q0:q0 + group], eps[q0:q0 + group],
                                                    relative_tolerance=rtol, precision=precision)
        return out

    legs = {"fp64, one per pass": ("fp64", 1), "fp64, 16 per pass": ("fp64_batched", 16),
            "fp64, 32 per pass": ("fp64_batched", 32), "mixed engine, 64 slots": ("mixed", 64)}
    solve("mixed", 64)  # engine allocation outside the timing
    results, rows = {}, {}
    for name, (precision, group) in legs.items():
        before = op.device.stats()
        t = time.perf_counter()  # each call returns host results: synchronous
        results[name] = solve(precision, group)
        wall = time.perf_counter() - t
        after = op.device.stats()
        rows[name] = {"solves_per_s": n_emissions / wall, "wall_s": round(wall, 3),
                      "reflections_or_sweeps": int(sum(r.iterations for r in results[name]))}
        s_v = 2 if after["integer_form"] else 8
        stream = after["nnz"] * (4 + s_v) + 8 * (after["n"] + 1)
        if precision == "fp64":
            passes, ms = after["reflections"] - before["reflections"], after["ms"] - before["ms"]
            moved = passes * (stream + 32 * after["n"])
        elif precision == "fp64_batched":
            passes, ms = after["batched_passes"] - before["batched_passes"], after["batched_ms"] - before["batched_ms"]
            slot_passes = after["batched_slot_passes"] - before["batched_slot_passes"]
            # per patch and slot: w gathered, eps, absorbed read + written, flight and w_next written
            moved = passes * (stream + 16 * after["n"]) + 48 * after["n"] * slot_passes
        else:
            continue
        rows[name].update({"passes": int(passes), "device_ms_per_pass": ms / passes if passes else None,
                           "achieved_GBps": moved / (ms * 1e-3) / 1e9 if ms > 0 else None})
    ref = results["fp64, one per pass"]
    for name, res in results.items():
        worst = 0.0
        for q in range(n_emissions):
            want = g.channels(ref[q].absorbed_power_W - E[q])
            got = g.channels(res[q].absorbed_power_W - E[q])
            scale = max(abs(want[0]), abs(want[1]))
            worst = max(worst, max(abs(a - b) for a, b in zip(got, want)) / scale)
        rows[name]["max_rel_flux_error"] = worst
    base = rows["fp64, one per pass"]["solves_per_s"]
    for row in rows.values():
        row["speedup"] = row["solves_per_s"] / base
    log("batching vs precision: " + ", ".join(f"{k} {v['solves_per_s']:.1f}/s" for k, v in rows.items()))
    return {"workload": f"{n_emissions} prescribed-field emissions of {g.n} patches, rtol {rtol:g}", "legs": rows}


BATCH_SIZES = (96, 128, 160, 192)
C5_K_SOLID = 1.0  # W/m/K: solid conductivity of the served coupled requests
# Jacobi-PCG steps of their Newton-Krylov solves: the default 10/20 leave FGMRES(30) short of its
# tolerance within 90 iterations at this conductivity ratio (k_s / k_g = 38)
C5_PCG_STEPS = (80,)


def batching_workload(sizes=BATCH_SIZES):
    """16 uneven geometries: foam, random fibres, aligned fibres and perforated plates at N = 96..192."""

    from porerad import synthetic

    return synthetic.mixed_geometries(sizes)


def small_batching(R, steps=5):
    """The same four families at N = 24..48: geometries too small to fill the chip one call at a time."""

    from porerad import _native, view_factors

    geos = [(name, view_factors.build_surface_patches(ph, 0.01)) for name, ph in batching_workload((24, 32, 40, 48))]
    seeds = [[10 + k] for k in range(len(geos))]
    n_rays = sum(P.n_patches for _, P in geos) * R
    batch = _native.TraceBatch([(P.geometry(), 0, P.n_patches, R, 1) for _, P in geos])
    batch.run(seeds)
    t = time.perf_counter()
    for _ in range(steps):
        batch.run(seeds)
    batched = (time.perf_counter() - t) / steps
    nnz = list(batch.nnz)
    batch.close()
    for k, (_, P) in enumerate(geos):
        view_factors.first_hit_edges(P, R, seeds[k]).close()
    t = time.perf_counter()
    for _ in range(steps):
        got = []
        for k, (_, P) in enumerate(geos):
            e = view_factors.first_hit_edges(P, R, seeds[k])
            got.append(e.nnz)
            e.close()
    calls = (time.perf_counter() - t) / steps
    if got != nnz:
        raise RuntimeError("batched and per-geometry edge counts differ")
    log(f"small geometries (N 24..48, {n_rays} rays): batched {batched * 1e3:.2f} ms, per-geometry calls "
        f"{calls * 1e3:.2f} ms ({calls / batched:.2f}x)")
    return {"geometries": {name: int(P.n_patches) for name, P in geos}, "rays": int(n_rays),
            "batched_ms": batched * 1e3, "per_geometry_calls_ms": calls * 1e3, "speedup": calls / batched,
            "rays_per_s_batched": n_rays / batched}


def run_batching(args, with_stream=True) -> dict:
    """One graph-captured trace pass over 16 uneven geometries (seeds 10..25, R rays per patch), replayed
    K times, against the same rays traced geometry by geometry; then the operators of all 16, assembled
    and balanced on the device from the pass's edge tables, serve a Poisson stream of radiosity and
    Newton-Krylov requests (run_mixed_stream)."""

    from porerad import _native, autotune, exchange, serving, view_factors

    R = args.rays
    t = time.perf_counter()
    geos = []
    for name, phases in batching_workload():
        geos.append((name, view_factors.build_surface_patches(phases, 0.01)))
    seeds = [[10 + k] for k in range(len(geos))]
    n_rays = sum(P.n_patches for _, P in geos) * R
    log(f"batch geometries ({time.perf_counter() - t:.1f} s): " +
        ", ".join(f"{name} {P.n_patches}" for name, P in geos) + f"; {n_rays} rays per pass")
    segments = [(P.geometry(), 0, P.n_patches, R, 1) for _, P in geos]
    tune, capacity, n_streams = None, 0, 0
    if args.tune:
        tune = autotune.tune_batch(segments, seeds, use_cache=False)
        capacity, n_streams = tune["ray_capacity"], tune["streams"]
        log("batch tuning: " + ", ".join(f"v{r['variant']}={r['graph_ms']:.1f}ms" for r in tune["table"]) +
            "; chunks x streams " + ", ".join(f"{r['ray_capacity']}x{r['streams']}={r['graph_ms']:.1f}ms"
                                             for r in tune["capacity_table"]))

    # plain path: one prh_trace_counts call per geometry
    seq_nnz = []
    largest = max(range(len(geos)), key=lambda k: geos[k][1].n_patches)
    view_factors.first_hit_edges(geos[largest][1], R, seeds[largest]).close()  # workspace sized
    t = time.perf_counter()
    for k, (_, P) in enumerate(geos):
        e = view_factors.first_hit_edges(P, R, seeds[k])
        seq_nnz.append(e.nnz)
        e.close()
    sequential_s = time.perf_counter() - t
    _native.trim_tracer()  # the per-call workspace of the largest geometry

    t = time.perf_counter()
    batch = _native.TraceBatch(segments, ray_capacity=capacity, streams=n_streams)
    create_s = time.perf_counter() - t
    log(f"trace batch: {batch.info} (created and captured in {create_s:.2f} s)")
    steps = max(1, min(args.steps, 3))
    for _ in range(max(1, min(args.warmup, 1))):
        batch.run(seeds)
    wall, dev = [], []
    for _ in range(steps):
        t = time.perf_counter()
        batch.run(seeds)
        wall.append(time.perf_counter() - t)
        dev.append(batch.elapsed_ms * 1e-3)
    step = float(np.mean(wall))
    if list(batch.nnz) != seq_nnz:
        raise RuntimeError("batched and per-geometry edge counts differ")
    reads = int(batch.occupancy_reads.sum())
    log(f"batched pass {step * 1e3:.1f} ms ({n_rays / step:.3e} rays/s, graph {np.mean(dev) * 1e3:.1f} ms); "
        f"per-geometry calls {sequential_s * 1e3:.1f} ms ({n_rays / sequential_s:.3e} rays/s)")
    algo_bytes = 4.0 * n_rays + float(reads)
    graph_s = float(np.mean(dev))
    record = {
        "workload": "foam, random fibre, aligned fibre, perforated plates x N in "
                    f"{list(BATCH_SIZES)}, seeds 10-25, {R} rays/patch, one replayed graph pass",
        "rays_per_s": n_rays / step,
        "replays": steps,
        "ms_per_replay": step * 1e3,
        "graph_ms": graph_s * 1e3,
        "per_geometry_calls_ms": sequential_s * 1e3,
        "speedup_vs_per_geometry_calls": sequential_s / step,
        "create_and_capture_s": create_s,
        "geometries": {name: int(P.n_patches) for name, P in geos},
        "rays_per_replay": int(n_rays),
        "edges": int(batch.nnz.sum()),
        "edge_counts_equal_per_geometry_calls": True,
        "layout": batch.info,
        "roofline": {"bound": "hbm", "achieved": algo_bytes / graph_s / 1e9, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": algo_bytes / graph_s / 1e9 / HBM_PEAK_GBPS, "traffic": None,
                     "kernel": "whole graph (generation + traversal + counting + compaction)",
                     "algorithmic_bytes_per_launch": algo_bytes, "occupancy_reads_per_ray": reads / n_rays},
    }
    if tune is not None:
        record["tuning"] = tune
    record["small_geometries"] = small_batching(R)
    if not with_stream:
        batch.close()
        return record

    # Each segment's edges move to a device table of their own (the batch's record buffers are then
    # freed) and each operator is assembled and balanced on the device from them.
    t = time.perf_counter()
    tables = [batch.edge_table(k) for k in range(len(geos))]
    batch.close()
    operators = {}
    for k, (name, P) in enumerate(geos):
        try:
            operators[name] = exchange.ExchangeOperator.from_edges(tables[k], np.full(P.n_patches, R, np.int64),
                                                                    P.patch_area_m2)
        finally:
            tables[k].close()
    setup = time.perf_counter() - t
    log(f"16 operators assembled and balanced on the device in {setup:.1f} s")
    record["stream"] = run_mixed_stream(dict(geos), operators)
    record["stream"]["operator_setup_s"] = round(setup, 1)
    for op in operators.values():
        op.device.close()
    return record


# Newton-Krylov requests of the served stream over the 16 geometries: (geometry, T_hot, emissivity, arrival
# index among the radiosity requests).
STREAM_COUPLED = (("foam96", 500.0, 0.3, 20), ("random_fibre96", 700.0, 0.6, 120),
                  ("aligned_fibre128", 900.0, 0.9, 220), ("foam128", 700.0, 0.6, 320))


def run_mixed_stream(patches, operators, n_requests=600, rate=200.0, seed=7):
    """A Poisson stream (rate 200/s, seed 7) of radiosity requests over the 16 resident operators
    (T_hot 500/700/900 K, eps 0.3/0.6/0.9, solver tolerance 1e-8) with Newton-Krylov coupled requests
    arriving among them (k_solid 1.0, k_gas 0.026 W/m/K, T_cold 300 K): served solves/s, p50/p99
    latency, and the coupled latencies, alone and in the stream."""

    from porerad import serving
    from porerad.constants import K_GAS

    def server():
        srv = serving.RadiosityServer()
        for name, op in operators.items():
            srv.add_geometry(name, patches[name], op)
        return srv

    rng = np.random.default_rng(seed)
    names = list(operators)
    arrivals = serving.poisson_arrivals(rate, n_requests, seed)
    reqs = [serving.RadiosityRequest(str(rng.choice(names)), float(rng.choice([500.0, 700.0, 900.0])), 300.0,
                                     float(rng.choice([0.3, 0.6, 0.9])), 1e-8, float(t), tag=i)
            for i, t in enumerate(arrivals)]
    creqs = [serving.CoupledRequest(g, Th, 300.0, eps, C5_K_SOLID, K_GAS, float(arrivals[at]), tag=f"c{j}",
                                    precision="mixed", preconditioner_steps=C5_PCG_STEPS)
             for j, (g, Th, eps, at) in enumerate(STREAM_COUPLED)]
    srv = server()
    srv.run([dataclasses.replace(r, arrival_s=0.0) for r in reqs[:: n_requests // 32]], real_time=False)  # warm
    t = time.perf_counter()
    replies = srv.run(reqs + creqs, real_time=True)
    wall = time.perf_counter() - t
    srv.close()
    rad = [r for r in replies if isinstance(r, serving.RadiosityReply)]
    cpl = [r for r in replies if isinstance(r, serving.CoupledReply)]
    rad_wall = max(r.finished_s for r in rad) - t
    out = {"workload": f"{n_requests} radiosity requests (rtol 1e-8) at {rate:.0f}/s Poisson (seed {seed}) over the "
                       f"16 geometries + {len(creqs)} Newton-Krylov requests (k_solid {C5_K_SOLID}, k_gas {K_GAS}, "
                       f"Jacobi-PCG {C5_PCG_STEPS[0]} steps, mixed-precision Jacobian solves)",
           "radiosity": serving.latency_summary(rad, rad_wall),
           "coupled": {"requests": [f"{g} T_hot {Th:.0f} K eps {eps}" for g, Th, eps, _ in STREAM_COUPLED],
                       "latency_s": {r.request.tag: round(r.latency_s, 3) for r in cpl},
                       "newton_updates": {r.request.tag: r.newton_updates for r in cpl},
                       "converged": all(r.converged for r in cpl)},
           "wall_s": wall}
    # each coupled request alone: latency, and the same result as in the stream
    alone = {}
    same = True
    by_tag = {r.request.tag: r for r in cpl}
    for r in creqs:
        srv = server()
        rep, = srv.run([dataclasses.replace(r, arrival_s=0.0)], real_time=False)
        srv.close()
        alone[r.tag] = round(rep.latency_s, 3)
        same &= rep.metrics == by_tag[r.tag].metrics
    out["coupled"]["alone_latency_s"] = alone
    out["coupled"]["alone_equals_in_stream"] = bool(same)
    log(f"stream over 16 geometries: radiosity {out['radiosity']['solves_per_s']:.1f} solves/s, p50 "
        f"{out['radiosity']['p50_ms']:.1f} ms, p99 {out['radiosity']['p99_ms']:.1f} ms; coupled latencies "
        f"{out['coupled']['latency_s']} s (alone {alone} s)")
    return out


def engine_roofline(before, after) -> dict:
    """HBM roofline of the served engine's ticks over a run.

    Algorithmic bytes of one tick with k running right-hand sides:
    nnz(S) (4 B column + s_v B value) + 8 (n + 1) row pointers + 128 n for
    the gathered FP16 vector of 64 slots read once + 24 n k for the running
    slots' own state (x and r read and written, rho read: 20 B in FP32; own
    z entry read and the next written: 4 B in FP16); the time is the ticks'
    device time from HIP events around each tick."""

    total_bytes = total_ms = ticks = 0.0
    for key, a in after.items():
        b = before.get(key, {"ticks": 0, "ms": 0.0, "slot_ticks": 0})
        n_t = a["ticks"] - b["ticks"]
        k_t = a["slot_ticks"] - b["slot_ticks"]
        total_bytes += (n_t * (a["nnz"] * (4 + a["value_bytes"]) + 8 * (a["n"] + 1) + 128 * a["n"])
                        + 24 * a["n"] * k_t)
        total_ms += a["ms"] - b["ms"]
        ticks += n_t
    achieved = total_bytes / (total_ms * 1e-3) / 1e9 if total_ms > 0 else 0.0
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBPS, "kernel": "k_tick", "ticks": int(ticks),
            "ms_per_tick": total_ms / ticks if ticks else None, "algorithmic_bytes": total_bytes}


def series_roofline(before, after) -> dict:
    """HBM roofline of the plain FP64 reflection series: per reflection
    nnz(S) (4 B column + s_v B value: 2 for the integer form K, 8 for FP64 S)
    + 8 (n + 1) + 32 n bytes (one FP64 vector gathered, four read or written),
    over the series' device time."""

    refl = after["reflections"] - before["reflections"]
    ms = after["ms"] - before["ms"]
    s_v = 2 if after["integer_form"] else 8
    total = refl * (after["nnz"] * (4 + s_v) + 8 * (after["n"] + 1) + 32 * after["n"])
    achieved = total / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBPS, "kernel": "k_series_step", "reflections": int(refl),
            "ms_per_reflection": ms / refl if refl else None}


def host_cpu() -> dict:
    """CPU model and core counts of this host."""

    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)),
            "omp_threads": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))}


def _omp_threads(n: int) -> None:
    import ctypes

    ctypes.CDLL("libgomp.so.1").omp_set_num_threads(int(n))


def _median_of(fn, runs=5):
    fn()  # warm-up
    times = []
    for _ in range(runs):
        t = time.perf_counter()
        fn()
        times.append(time.perf_counter() - t)
    return float(np.median(times)), times


def cpu_baseline(phases, R, seed, target_s=2.0):
    """The C oracle (first_hit_counts of a fixed patch range: generation, FP64 DDA, sorted
    counts) on all host threads and on one: one warm-up, then the median of 5 runs."""

    from oracle import oracle as O

    Q = O.build_surface_patches(phases, 0.01)
    cpu = host_cpu()
    threads = cpu["omp_threads"]

    def sample(seconds):
        t = time.perf_counter()
        O.first_hit_counts_range(Q, R, seed, 0, 1000)
        rate = 1000 / max(time.perf_counter() - t, 1e-6)  # patches per second
        return int(min(Q.n_patches, max(200, rate * seconds)))

    out = {"unit": "rays/s", "kind": "port", "cpu": cpu}
    _omp_threads(threads)
    n_all = sample(target_s)
    med, runs = _median_of(lambda: O.first_hit_counts_range(Q, R, seed, 0, n_all))
    out.update({"value": n_all * R / med, "cores": threads, "median_s": med, "runs_s": runs,
                "sample": f"C oracle first_hit_counts of foam patches [0, {n_all}) x {R} rays (seed {seed}), "
                          f"{threads} OpenMP threads, 1 warm-up + median of 5"})
    _omp_threads(1)
    try:
        n_one = sample(target_s)
        med1, runs1 = _median_of(lambda: O.first_hit_counts_range(Q, R, seed, 0, n_one))
    finally:
        _omp_threads(threads)
    out["one_core"] = {"value": n_one * R / med1, "median_s": med1, "runs_s": runs1,
                       "sample": f"patches [0, {n_one}) x {R} rays, 1 thread, 1 warm-up + median of 5"}
    return out


def cpu_radiosity_baseline(P, op, T_hot=700.0, eps=0.6, rtol=1e-8):
    """FP64 reflection series of the C oracle (exchange.py:217-245 on the operator's own symmetric CSR,
    row-parallel over all host threads) for one prescribed-field request: solves/s from one warm-up
    and the median of 3 solves."""

    from oracle import oracle as O
    from porerad import coupled, exchange

    rowptr, col, val = op.device.csr()
    d = np.asarray(op.scale, np.float64)
    rows = np.repeat(np.arange(P.n_patches, dtype=np.int64), np.diff(rowptr))
    G = val * d[rows] * d[col]  # diag(d) S diag(d): the map of rmatvec, x -> G (x / A)
    del rows, val
    T_m = 0.5 * (T_hot + 300.0)
    T = coupled.patch_temperature(P, coupled.linear_cell_temperature(P.solid.shape, T_m, T_hot - 300.0),
                                  T_hot, 300.0)
    e = np.full(P.n_patches, eps)
    E = exchange.emission(e, P.patch_area_m2, T, T_m)
    n = P.n_patches
    absorbed, flight, work = np.empty(n), np.empty(n), np.empty(n)
    residual = np.zeros(1)
    area = np.ascontiguousarray(P.patch_area_m2, np.float64)
    state = {}

    def solve():
        state["refl"] = O.lib().orc_reflection_series_csr(
            O._p(rowptr), O._p(col), O._p(G), n, O._p(area), O._p(E), O._p(e), rtol, 10000, O._p(absorbed),
            O._p(flight), O._p(work), O._p(residual))

    med, runs = _median_of(solve, runs=3)
    threads = host_cpu()["omp_threads"]
    return {"value": 1.0 / med, "unit": "solves/s", "cores": threads, "kind": "port", "median_s": med,
            "runs_s": runs, "reflections": int(state["refl"]),
            "sample": f"C oracle FP64 reflection series (rtol {rtol:g}) on the foam operator's symmetric CSR "
                      f"({op.device.nnz_symmetric} nonzeros), T_hot {T_hot:.0f} K, eps {eps}, {threads} OpenMP "
                      "threads, 1 warm-up + median of 3"}


def run_sweep_probes(R, reps=20) -> dict:
    """What one batched sweep of the foam operator costs per slot format, and
    what dense MFMA tiles would take off it.

    Formats: FP32 x 32 slots, FP16 x 64 (the engine's) and FP64 x 16 (the
    same 128 B per gathered column), FP16 x 32 (64 B) and FP64 x 32 (256 B);
    the slot-sweep rate is slots / ms.  Tiles: 16 x 4 blocks of K holding at
    least 4, 8 or 16 nonzeros multiplied by v_mfma_f32_16x16x4f32, the rest
    gathered; the plain gather's time against tiles + remainder."""

    from porerad import _native, view_factors

    P = view_factors.build_surface_patches(foam_workload(), 0.01)
    op = build_operator(P, R, 3)
    slots = {"fp32x32": 32, "fp16x64": 64, "fp16x32": 32, "fp64x16": 16, "fp64x32": 32}
    formats = {}
    for name, fmt in _native.PROBE_FORMATS.items():
        ms = op.device.probe_gather(fmt, reps)
        formats[name] = {"ms_per_sweep": ms, "slot_sweeps_per_ms": slots[name] / ms}
        log(f"sweep {name}: {ms:.3f} ms, {slots[name] / ms:.1f} slot-sweeps/ms")
    base = formats["fp32x32"]["slot_sweeps_per_ms"]
    for row in formats.values():
        row["vs_fp32x32"] = row["slot_sweeps_per_ms"] / base
    orders = {}
    for name, order in patch_orders(P).items():
        o = op.device.probe_order(order, reps)
        orders[name] = o
        log(f"patch order {name}: stored {o['stored_ms']:.3f} ms (streamed matrix {o['stored_streamed_ms']:.3f}), "
            f"reordered {o['reordered_ms']:.3f} ms (streamed {o['reordered_streamed_ms']:.3f})")
    panels = {}
    for k in (4, 8, 16):
        pn = op.device.probe_panels(k, reps)
        panels[f"panels_{k}"] = pn
        log(f"column panels {k}: plain {pn['plain_ms']:.3f} ms, panels {pn['panels_ms']:.3f} ms "
            f"(sweep order {pn['panels_sweep_order_ms']:.3f} ms)")
    tiles = {}
    for fill in (4, 8, 16):
        t = op.device.probe_tiles(fill, reps)
        t["tiles_plus_remainder_ms"] = t["tiles_ms"] + t["remainder_ms"]
        t["speedup_vs_plain"] = t["plain_ms"] / t["tiles_plus_remainder_ms"]
        tiles[f"min_fill_{fill}"] = t
        log(f"tiles >= {fill}: {t['tiles']} tiles hold {t['nnz_fraction_in_tiles']:.3f} of nnz; plain "
            f"{t['plain_ms']:.3f} ms vs tiles {t['tiles_ms']:.3f} + remainder {t['remainder_ms']:.3f} ms")
    op.device.close()
    return {"workload": f"foam N=128, {P.n_patches} patches, {R} rays/patch (seed 3)", "reps": reps,
            "formats": formats, "patch_orders": orders, "column_panels": panels, "mfma_tiles": tiles}


def patch_orders(P) -> dict:
    """Renumberings of the patches for the locality probe: Morton (Z-order) and
    z-major lexicographic order of the patch origins on the half-voxel grid."""

    ijk = np.floor(np.asarray(P.patch_origin) * 2.0).astype(np.int64)
    out = {"morton": P.sweep_order(), "hilbert": P.sweep_order("hilbert"),
           "z-major": np.lexsort((ijk[:, 0], ijk[:, 1], ijk[:, 2])).astype(np.int32)}
    return out


def run_monte_carlo(phases, with_cpu: bool) -> dict:
    """Multi-bounce Monte Carlo statistics of the foam (properties.py / penetration.py):
    identify_properties (2 seeds x 200 000 beam rays) and penetration_response
    (2 seeds x 500 000 Lambertian rays) at eps 0.6 along x, timed on the device after
    one warm call (the cos/sin table is built then); the CPU oracle on the same rays."""

    from porerad import penetration, properties

    eps = 0.6
    properties.identify_properties(phases, "x", eps, n_rays=1000)  # table and kernels warm
    t = time.perf_counter()
    props = properties.identify_properties(phases, "x", eps)
    t_props = time.perf_counter() - t
    t = time.perf_counter()
    pen = penetration.penetration_response(phases, "x", eps)
    t_pen = time.perf_counter() - t
    out = {"workload": "voronoi foam N=128^3, axis x, eps 0.6, 'article' restart",
           "properties": {"rays": props["n_rays"], "s": t_props, "rays_per_s": props["n_rays"] / t_props,
                          "beta": props["beta"], "omega": props["omega"], "g": props["g"]},
           "penetration": {"rays": pen["n_rays"], "s": t_pen, "rays_per_s": pen["n_rays"] / t_pen,
                           "reflectance": pen["reflectance"], "transmittance": pen["transmittance"],
                           "absorptance": pen["absorptance"]}}
    if with_cpu:
        from oracle import oracle as O

        solid = (np.asarray(phases) == 1).astype(np.uint8)
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        t = time.perf_counter()
        for seed in penetration.SEEDS:
            O.mc_penetration(solid, penetration.N_RAYS, seed, eps)
        t_cpu = time.perf_counter() - t
        out["cpu_penetration"] = {"rays_per_s": 2 * penetration.N_RAYS / t_cpu, "cores": threads, "kind": "port",
                                  "s": t_cpu}
    return out


def run_microstructure(phases, with_cpu: bool) -> dict:
    """The step before the path (geometry.py / descriptors.py): one core microstructure of
    each structure at N = 128, phi_s 0.10 (OF seed 0: 32 nodes and 59 struts, three bisection
    renders; FA seed 2: long-fibre scatter; TL seed 0: layered field threshold) built on
    the device, and the scalar plus directional descriptors of the foam, each timed after
    one warm call; the CPU oracle (C stamping on every host thread, NumPy / SciPy
    descriptors) on the same work."""

    from porerad import descriptors, geometry

    builds = (("OF", 0), ("FA", 2), ("TL", 0))
    geometry.build("FI", 32, 0.10, 0)  # kernels warm
    out = {"workload": "core microstructures N=128^3 phi_s 0.10 (OF s0, FA s2, TL s0); descriptors of the foam"}
    for s, seed in builds:
        t = time.perf_counter()
        p = geometry.build(s, 128, 0.10, seed)
        out[f"build_{s}_s"] = time.perf_counter() - t
        out[f"build_{s}_phi"] = float(np.mean(p == 1))
    descriptors.scalar_descriptors(phases)
    t = time.perf_counter()
    descriptors.scalar_descriptors(phases)
    out["scalar_descriptors_s"] = time.perf_counter() - t
    t = time.perf_counter()
    descriptors.directional_descriptors(phases)
    out["directional_descriptors_s"] = time.perf_counter() - t
    if with_cpu:
        from oracle import descriptors_oracle as DO
        from oracle import geometry_oracle as GO

        cpu = {"cores": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)), "kind": "port"}
        for s, seed in builds:
            t = time.perf_counter()
            GO.build(s, 128, 0.10, seed)
            cpu[f"build_{s}_s"] = time.perf_counter() - t
        t = time.perf_counter()
        DO.scalar_descriptors(phases)
        cpu["scalar_descriptors_s"] = time.perf_counter() - t
        t = time.perf_counter()
        DO.directional_descriptors(phases)
        cpu["directional_descriptors_s"] = time.perf_counter() - t
        out["cpu"] = cpu
    return out


def copy_probe(n_bytes: int = 1 << 30) -> None:
    """One plain device copy of n_bytes (a torch elementwise copy kernel): the
    reference point for scaling the FETCH_SIZE / WRITE_SIZE counters."""

    import torch

    src = torch.ones(n_bytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    log(f"copy probe: {n_bytes} bytes read, {n_bytes} bytes written")


def main() -> None:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--rays", type=int, default=1000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-radiosity", action="store_true", help="skip the radiosity solves")
    ap.add_argument("--tune", action="store_true", help="time every tracer variant first and use the fastest")
    ap.add_argument("--copy-probe", action="store_true", help="only run one plain 1 GiB device copy")
    ap.add_argument("--microstructure", action="store_true",
                    help="only the microstructure leg (generators and descriptors)")
    ap.add_argument("--monte-carlo", action="store_true", help="only the Monte Carlo leg")
    ap.add_argument("--sweep-probes", action="store_true",
                    help="only the sweep probes: slot formats and MFMA tiles on the foam operator")
    ap.add_argument("--batching", action="store_true",
                    help="the 16-geometry batched trace pass (+ 64 radiosity RHS per geometry) instead")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe()
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU; --gpus above 1 is not supported "
              "(the tracer would shard by patch range, the radiosity by request)", file=sys.stderr)
        sys.exit(2)

    from porerad import _native, view_factors

    log(_native.lib().prh_version().decode())
    if args.batching:
        print(json.dumps(run_batching(args)), flush=True)
        return
    if args.microstructure:
        print(json.dumps(run_microstructure(foam_workload(), not args.no_cpu)), flush=True)
        return
    if args.monte_carlo:
        print(json.dumps(run_monte_carlo(foam_workload(), not args.no_cpu)), flush=True)
        return
    if args.sweep_probes:
        print(json.dumps(run_sweep_probes(args.rays)), flush=True)
        return
    t = time.perf_counter()
    phases = foam_workload()
    P = view_factors.build_surface_patches(phases, 0.01)
    log(f"foam N=128 phi={phases.mean():.4f}: {P.n_patches} patches ({time.perf_counter() - t:.1f} s)")
    R, seed = args.rays, 3
    n_rays = P.n_patches * R

    tune = None
    if args.tune:
        from porerad import autotune

        tune = autotune.tune_tracer(P, R, [seed], use_cache=False)
        log("tracer variants: " + ", ".join(f"v{r['variant']}(refill {r['refill_min']}, occupancy {r['occupancy']})="
                                            f"{r['trace_ms']:.2f}ms" for r in tune["table"]))
    wall, parts, reads, nnz = run_tracer(P, R, seed, args.steps, args.warmup)
    step = float(np.mean(wall))
    gen_s, trace_s, count_s, compact_s = (float(x) for x in parts.mean(axis=0))
    log(f"trace step {step * 1e3:.2f} ms (generation {gen_s * 1e3:.2f}, traversal {trace_s * 1e3:.2f}, "
        f"counting {count_s * 1e3:.2f}, compaction {compact_s * 1e3:.2f} ms), {nnz} edges, "
        f"{reads / n_rays:.1f} occupancy reads/ray")
    parity = count_parity(P, R, seed)
    log(f"counts equal to the CPU oracle's (count_sha256 {parity.get('count_sha256', '-')[:16]}...)"
        if parity["checked"] else "count parity not checked: " + parity["reason"])
    kernel = trace_s
    kernel_name = traversal_kernel()
    lanes = lane_share(P, R, seed, reads) if kernel_name == "k_trace_lean" else None
    algo_bytes = 4.0 * n_rays + float(reads)
    achieved = algo_bytes / kernel / 1e9
    record = {
        "metric": METRIC,
        "value": n_rays / step,
        "unit": "rays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": step * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "voronoi foam N=128^3, 64 seeds, phi=0.10 (seed 2), axis x, "
                               f"{R} rays/patch (seed 3)",
                   "n_patches": int(P.n_patches), "rays_per_step": int(n_rays), "edges": int(nnz)},
        "parity": parity,
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBPS,
            # FETCH_SIZE x 2 + WRITE_SIZE x 1 (copy-calibrated) of this kernel at
            # the default workload, from the rocprofv3 --pmc passes summarised in
            # traffic_source; null for other ray counts
            "traffic": measured_traffic(kernel_name) if R == 1000 else None,
            "traffic_source": os.path.relpath(traffic_profile(kernel_name), REPO)
            + " (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE)",
            "kernel": kernel_name,
            # the kernel's actual limit: VALU issue (SQ_INSTS_VALU x 4 cycles over
            # the SIMDs' busy cycles) at the default workload, from the
            # rocprofv3 --pmc pass summarised in valu_issue_source
            "valu_issue_frac": profile_value(counters_profile(kernel_name), "valu_issue_frac") if R == 1000 else None,
            "valu_issue_source": os.path.relpath(counters_profile(kernel_name), REPO)
            + " (rocprofv3 --pmc SQ_INSTS_VALU, SQ_BUSY_CYCLES)",
            "lane_share": lanes,
            "kernel_ms": kernel * 1e3,
            "algorithmic_bytes_per_launch": algo_bytes,
            "occupancy_reads_per_ray": reads / n_rays,
            "dda_steps_per_s": float(reads) / kernel,
            "stage_ms": {"generation": gen_s * 1e3, "traversal": trace_s * 1e3, "counting": count_s * 1e3,
                         "compaction": compact_s * 1e3},
        },
    }
    if tune is not None:
        record["tracer_variant"] = {"chosen": tune["variant"], "table": tune["table"]}
    if not args.no_radiosity:
        rad = run_radiosity(P, R, seed, with_cpu=not args.no_cpu)
        record["radiosity"] = rad
        log(f"radiosity plain fp64 {rad['plain_fp64']['solves_per_s']:.1f} solves/s; served "
            f"{rad['served']['solves_per_s']:.1f} solves/s at {rad['served']['arrival_rate_per_s']:.0f}/s offered, "
            f"p50 {rad['served']['p50_ms']:.1f} ms p99 {rad['served']['p99_ms']:.1f} ms; saturation "
            f"{rad['saturation']['solves_per_s']:.1f} solves/s ({rad['speedup_vs_plain_fp64']:.1f}x plain); "
            f"max flux error {max(rad['max_rel_flux_error_vs_fp64'].values()):.1e}")
        rf = rad["roofline"]
        log(f"radiosity roofline: engine ticks {rf['engine']['achieved']:.0f} GB/s ({rf['engine']['frac'] * 100:.1f} % "
            f"of HBM, {rf['engine']['ms_per_tick']:.3f} ms/tick); FP64 series {rf['fp64_series']['achieved']:.0f} GB/s "
            f"({rf['fp64_series']['frac'] * 100:.1f} %, {rf['fp64_series']['ms_per_reflection']:.3f} ms/reflection)")
        cb = rad["serving_with_coupled"]
        log(f"precision at 1e-10: {rad['precision']['max_rel_error_per_flux']}")
        log(f"with coupled solves: radiosity {cb['radiosity']['solves_per_s']:.1f} solves/s p50 "
            f"{cb['radiosity']['p50_ms']:.1f} ms p99 {cb['radiosity']['p99_ms']:.1f} ms; coupled latencies "
            f"{cb['coupled']['latency_s']} s (alone: {cb['coupled']['alone_latency_s']} s), "
            f"Newton updates {cb['coupled']['newton_updates']}, max flux error vs FP64 "
            f"{max(max(v.values()) for v in cb['coupled']['max_rel_flux_error_vs_fp64'].values()):.2e}")
    if not args.no_radiosity:
        mc = run_monte_carlo(phases, not args.no_cpu)
        record["monte_carlo"] = mc
        log(f"monte carlo: properties {mc['properties']['rays_per_s']:.3e} rays/s, penetration "
            f"{mc['penetration']['rays_per_s']:.3e} rays/s"
            + (f" (CPU oracle {mc['cpu_penetration']['rays_per_s']:.3e} rays/s on "
               f"{mc['cpu_penetration']['cores']} threads)" if "cpu_penetration" in mc else ""))
    if not args.no_radiosity:
        ms = run_microstructure(phases, not args.no_cpu)
        record["microstructure"] = ms
        log("microstructure: build OF {build_OF_s:.3f} s, FA {build_FA_s:.3f} s, TL {build_TL_s:.3f} s; descriptors "
            "scalar {scalar_descriptors_s:.4f} s, directional {directional_descriptors_s:.4f} s".format(**ms)
            + (" (CPU oracle: OF {build_OF_s:.2f} s, FA {build_FA_s:.2f} s, TL {build_TL_s:.2f} s, scalar "
               "{scalar_descriptors_s:.3f} s, directional {directional_descriptors_s:.3f} s)".format(**ms["cpu"])
               if "cpu" in ms else ""))
    if not args.no_radiosity:
        _native.trim_tracer()
        record["batching"] = run_batching(args)
    if not args.no_cpu:
        record["cpu_baseline"] = cpu_baseline(phases, R, seed)
        cb = record["cpu_baseline"]
        log(f"cpu baseline {cb['value']:.3e} rays/s on {cb['cores']} threads ({cb['one_core']['value']:.3e} on one "
            f"core), {cb['cpu']['model']}")
    print(json.dumps(record), flush=True)


if __name__ == "__main__":
    main()
