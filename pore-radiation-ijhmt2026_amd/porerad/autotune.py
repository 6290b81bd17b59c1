"""Launch-shape autotuner of the HIP tracer.

The traversal kernel comes in a few compiled variants (refill threshold x
occupancy test, see ``prh_tracer_variants``).  Which one is fastest depends
on the geometry (mean ray length, divergence of ray lengths, how much of the
pore space the LDS brick bitset clears) and on the ray
count per patch.  ``tune_tracer`` times every variant on a sample of the
patches, keeps the fastest and records the choice in a small JSON cache
keyed by the workload signature, so later runs start from it.

``tune_batch`` does the same for a multi-geometry trace batch and also
picks its chunk size (ray records per chunk of the captured graph).

Results never depend on the variant: every variant produces bit-identical
counts (checked by tests/test_gpu_tracer.py).
"""

from __future__ import annotations

import json
import os
import time

from . import _native

CACHE_ENV = "PORERAD_TUNE_CACHE"


def _cache_path() -> str:
    return os.environ.get(CACHE_ENV) or os.path.join(os.path.expanduser("~"), ".cache", "porerad", "tracer_tune.json")


def _load() -> dict:
    try:
        with open(_cache_path()) as handle:
            return json.load(handle)
    except (OSError, ValueError):
        return {}


def _store(cache: dict) -> None:
    path = _cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        with open(path, "w") as handle:
            json.dump(cache, handle, indent=1)
    except OSError:
        pass


def signature(patches, rays_per_patch: int, n_seeds: int) -> str:
    nx, ny, nz = patches.solid.shape
    return f"{nx}x{ny}x{nz}/p{patches.n_patches}/r{int(rays_per_patch) * int(n_seeds)}"


def tune_tracer(patches, rays_per_patch: int, seeds, *, sample_patches: int = 40000, repeats: int = 2,
                use_cache: bool = True) -> dict:
    """Pick the fastest traversal variant for this workload; returns the timing table."""

    from .view_factors import first_hit_edges

    seeds = [int(s) for s in seeds]
    key = signature(patches, rays_per_patch, len(seeds))
    cache = _load() if use_cache else {}
    if key in cache:
        _native.set_tracer_variant(cache[key]["variant"])
        return cache[key]
    n = patches.n_patches
    p0 = max(0, n // 2 - sample_patches // 2)
    p1 = min(n, p0 + sample_patches)
    table = []
    for v in _native.tracer_variants():
        _native.set_tracer_variant(v["variant"])
        first_hit_edges(patches, rays_per_patch, seeds, patch_range=(p0, p1)).close()
        best = float("inf")
        for _ in range(repeats):
            t = time.perf_counter()
            first_hit_edges(patches, rays_per_patch, seeds, patch_range=(p0, p1)).close()
            best = min(best, _native.last_timings()[1])
            _ = time.perf_counter() - t
        table.append({"variant": v["variant"], "refill_min": v["refill_min"], "occupancy": v["occupancy"],
                      "trace_ms": best})
    winner = min(table, key=lambda row: row["trace_ms"])
    _native.set_tracer_variant(winner["variant"])
    result = {"variant": winner["variant"], "table": table, "sample": [p0, p1]}
    if use_cache:
        cache[key] = result
        _store(cache)
    return result


def batch_signature(segments, n_seeds) -> str:
    parts = [f"{'x'.join(str(s) for s in g.shape)}:{p1 - p0}:{int(r) * int(s)}"
             for (g, p0, p1, r, _), s in zip(segments, n_seeds)]
    return "batch/" + "/".join(parts)


def tune_batch(segments, seeds, *, sample_fraction: float = 0.125, capacities=(1 << 26, 1 << 27, 1 << 28),
               streams=(1, 2, 3), use_cache: bool = True) -> dict:
    """Pick the traversal variant, chunk size and stream count of a multi-geometry trace batch.

    ``segments`` are ``TraceBatch`` segments ``(geometry, p0, p1, rays_per_patch,
    n_seeds)`` and ``seeds`` their seed lists.  The variant is timed on the
    middle ``sample_fraction`` of every segment's patch range, then every
    (chunk size in ray records, streams) pair on the whole batch with that
    variant.  Times are the device time of one replay after a warm replay.
    Leaves the chosen variant set; returns the tables.
    """

    key = batch_signature(segments, [len(s) for s in seeds])
    cache = _load() if use_cache else {}
    if key in cache:
        _native.set_tracer_variant(cache[key]["variant"])
        return cache[key]

    def timed(segs, capacity=0, n_streams=0):
        batch = _native.TraceBatch(segs, ray_capacity=capacity, streams=n_streams)
        try:
            batch.run(seeds)
            batch.run(seeds)
            return batch.elapsed_ms, batch.info
        finally:
            batch.close()

    sample = []
    for g, p0, p1, r, s in segments:
        span = p1 - p0
        m = max(1, int(span * sample_fraction)) if span else 0
        a = p0 + (span - m) // 2
        sample.append((g, a, a + m, r, s))
    table = []
    for v in _native.tracer_variants():
        _native.set_tracer_variant(v["variant"])
        ms, _ = timed(sample)
        table.append({"variant": v["variant"], "refill_min": v["refill_min"], "occupancy": v["occupancy"],
                      "graph_ms": ms})
    winner = min(table, key=lambda row: row["graph_ms"])
    _native.set_tracer_variant(winner["variant"])
    cap_table = []
    for cap in capacities:
        for n_streams in streams:
            ms, info = timed(segments, cap, n_streams)
            cap_table.append({"ray_capacity": int(cap), "streams": info["streams"], "chunks": info["chunks"],
                              "graph_ms": ms, "device_bytes": info["device_bytes"]})
    best = min(cap_table, key=lambda row: row["graph_ms"])
    result = {"variant": winner["variant"], "ray_capacity": best["ray_capacity"], "streams": best["streams"],
              "table": table, "capacity_table": cap_table, "sample_fraction": sample_fraction}
    if use_cache:
        cache[key] = result
        _store(cache)
    return result
