"""Benchmark of the porerad radiation hot path on one MI355X.

Workload (one GPU): open-cell foam N=128^3 from the periodic Voronoi edges of
64 uniform seeds, strut radius bisected to solid fraction 0.10 (seed 2), axis
x, 1000 cosine-weighted rays per patch (seed 3).

A step is one fused first-hit pass over every patch of the foam: seeded ray
generation -> FP64 DDA -> per-patch integer counts -> sorted edge table in
HBM (prh_trace_counts).  ``value`` is first-hit rays per second of the whole
step with the geometry already resident.  Alongside it the line reports the
served radiosity solves (FP64 reflection series to 1e-8 per request on the
device-resident exchange operator of the same foam), the roofline of the
trace kernel and the CPU oracle timed on this host.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""

from __future__ import annotations

import argparse
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


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def foam_workload():
    from porerad import synthetic

    return synthetic.voronoi_foam(128, 64, 0.10, 2)


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


def run_radiosity(P, R, seed, eps_list, repeats):
    from porerad import coupled, exchange, view_factors

    t0 = time.perf_counter()
    o, r, c = view_factors.first_hit_counts_pooled(P, R, [seed])
    op = exchange.ExchangeOperator(o, r, c, np.full(P.n_patches, R, np.int64), P.patch_area_m2)
    setup = time.perf_counter() - t0
    log(f"operator: {op.device.nnz_symmetric} symmetric nonzeros, balance {op.balance_method} "
        f"({op.balance_iterations}), built in {setup:.1f} s")
    nx = P.solid.shape[0]
    T_h, T_c = 700.0, 300.0
    lat, iters = [], []
    for _ in range(repeats):
        for eps in eps_list:
            t = time.perf_counter()
            state = coupled.evaluate_radiation(op, P, coupled.linear_cell_temperature(P.solid.shape, 500.0, 400.0),
                                               T_h, T_c, eps, relative_tolerance=1e-8,
                                               reference_temperature_K=500.0)
            lat.append(time.perf_counter() - t)
            iters.append(state.radiosity.iterations)
    lat = np.array(lat)
    return {
        "solves": int(lat.size),
        "solves_per_s": float(lat.size / lat.sum()),
        "p50_ms": float(np.percentile(lat, 50) * 1e3),
        "p99_ms": float(np.percentile(lat, 99) * 1e3),
        "precision": "fp64 reflection series, rtol 1e-8",
        "emissivities": list(eps_list),
        "reflections": [int(i) for i in iters[: len(eps_list)]],
        "nnz_symmetric": int(op.device.nnz_symmetric),
        "operator_build_s": round(setup, 2),
        "n_patches": int(P.n_patches),
        "nx": int(nx),
    }


def cpu_baseline(phases, R, seed, budget_s=12.0):
    from oracle import oracle as O

    Q = O.build_surface_patches(phases, 0.01)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    n_p = 200
    start = 0
    done_rays, spent = 0, 0.0
    while spent < budget_s:
        n_p = min(n_p, Q.n_patches - start)
        t = time.perf_counter()
        O.first_hit_counts_range(Q, R, seed, start, start + n_p)
        dt = time.perf_counter() - t
        done_rays += n_p * R
        spent += dt
        start = (start + n_p) % Q.n_patches
        n_p = int(min(max(n_p, n_p * 0.5 * (budget_s - spent) / max(dt, 1e-3)), 50000))
    return {
        "value": float(done_rays / spent),
        "unit": "rays/s",
        "cores": threads,
        "kind": "port",
        "sample": f"C oracle (OpenMP, {threads} threads) on {done_rays // R} foam patches x {R} rays "
                  f"({spent:.1f} s of CPU work)",
    }


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
    ap.add_argument("--no-tune", action="store_true", help="keep the default tracer variant")
    ap.add_argument("--copy-probe", action="store_true", help="only run one plain 1 GiB device copy")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe()
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU; multi-GPU runs are not part of this build "
              "(the tracer shards by patch range, the radiosity by request)", file=sys.stderr)
        sys.exit(2)

    from porerad import _native, view_factors

    log(_native.lib().prh_version().decode())
    t = time.perf_counter()
    phases = foam_workload()
    P = view_factors.build_surface_patches(phases, 0.01)
    log(f"foam N=128 phi={phases.mean():.4f}: {P.n_patches} patches ({time.perf_counter() - t:.1f} s)")
    R, seed = args.rays, 3
    n_rays = P.n_patches * R

    tune = None
    if not args.no_tune:
        from porerad import autotune

        tune = autotune.tune_tracer(P, R, [seed], use_cache=False)
        log("tracer variants: " + ", ".join(f"v{r['variant']}(refill {r['refill_min']}, coarse {r['coarse']})="
                                            f"{r['trace_ms']:.2f}ms" for r in tune["table"]))
    wall, parts, reads, nnz = run_tracer(P, R, seed, args.steps, args.warmup)
    step = float(np.mean(wall))
    gen_s, trace_s, count_s, compact_s = (float(x) for x in parts.mean(axis=0))
    log(f"trace step {step * 1e3:.2f} ms (generation {gen_s * 1e3:.2f}, traversal {trace_s * 1e3:.2f}, "
        f"counting {count_s * 1e3:.2f}, compaction {compact_s * 1e3:.2f} ms), {nnz} edges, "
        f"{reads / n_rays:.1f} occupancy reads/ray")
    kernel = trace_s
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
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBPS,
            "traffic": None,
            "kernel": "k_trace_stream",
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
        record["radiosity"] = run_radiosity(P, R, seed, [0.3, 0.6, 0.9], 2)
        log(f"radiosity: {record['radiosity']['solves_per_s']:.1f} solves/s, p50 "
            f"{record['radiosity']['p50_ms']:.2f} ms, p99 {record['radiosity']['p99_ms']:.2f} ms")
    if not args.no_cpu:
        record["cpu_baseline"] = cpu_baseline(phases, R, seed)
        log(f"cpu baseline {record['cpu_baseline']['value']:.3e} rays/s")
    print(json.dumps(record), flush=True)


if __name__ == "__main__":
    main()
