"""Multi-geometry trace batches (prh_trace_batch_*) on the GPU.

Each segment of a batch must give, bit for bit, the pooled edge table and
occupancy-read count of a single prh_trace_counts call on the same geometry,
patch range, rays per patch and seeds -- whatever the chunking, and on every
replay of the captured graph.  The single-geometry path is pinned against the
CPU oracle and the reference's archived hashes (test_gpu_tracer.py); here one
case is also checked against the oracle directly, and one batch of all the
archived hash cases reproduces every archived SHA-256.
"""

from __future__ import annotations

import numpy as np
import pytest

from golden_data import cases, phases_x
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    from porerad import _native, view_factors

    _native.lib()
    return _native, view_factors


def _random(seed, shape, density):
    rng = np.random.default_rng(seed)
    return (rng.random(shape) < density).astype(np.uint8)


def _plate(shape, lo, hi):
    phases = np.zeros(shape, np.uint8)
    phases[lo:hi] = 1
    phases[lo:hi, 1:3, 2:4] = 0  # one hole
    return phases


@pytest.fixture(scope="module")
def media(env):
    _, vf = env
    return [
        vf.build_surface_patches(_random(1, (20, 16, 16), 0.15), 0.01),
        vf.build_surface_patches(_random(2, (12, 9, 11), 0.25), 0.004),
        vf.build_surface_patches(_plate((14, 8, 8), 6, 8), 0.01),
        vf.build_surface_patches(np.zeros((6, 5, 7), np.uint8), 0.006),
        vf.build_surface_patches(_random(3, (28, 24, 22), 0.1), 0.02),
    ]


def _segments(media):
    """(patches, p0, p1, rays per patch, seeds) of uneven segments."""

    a, b, c, d, e = media
    return [
        (a, 0, a.n_patches, 37, [5]),
        (b, 0, b.n_patches, 64, [7, 8]),
        (c, 0, c.n_patches, 1000, [9]),
        (a, a.n_patches // 3, a.n_patches // 3 + 101, 3, [1, 2, 3]),
        (d, 0, d.n_patches, 50, [4]),
        (e, 0, e.n_patches, 200, [11]),
        (c, 2, 2, 16, [6]),  # empty range
        (b, 0, b.n_patches, 2500, [12, 13]),  # more than 1024 rays per patch
    ]


def _single(vf, P, p0, p1, R, seeds):
    edges = vf.first_hit_edges(P, R, seeds, patch_range=(p0, p1))
    try:
        return edges.to_host(), edges.occupancy_reads
    finally:
        edges.close()


@pytest.mark.parametrize("capacity,streams,graph", [(0, 0, True), (1 << 16, 1, True), (3 << 14, 3, True),
                                                    (1 << 16, 2, False), (1 << 15, 4, True)])
def test_batch_equals_single_calls_on_every_replay(env, media, capacity, streams, graph):
    native, vf = env
    segs = _segments(media)
    batch = native.TraceBatch([(P.geometry(), p0, p1, R, len(s)) for P, p0, p1, R, s in segs], ray_capacity=capacity,
                              streams=streams, graph=graph)
    try:
        assert batch.info["graph"] == int(graph)
        if capacity:
            assert batch.info["chunks"] > 1 and batch.info["pieces"] > len(segs)
            assert batch.info["streams"] == (streams or 2)
        assert batch.info["rays"] == sum((p1 - p0) * R * len(s) for _, p0, p1, R, s in segs)
        reseeded = [[x + 1000 for x in s] for *_, s in segs]
        for seeds in ([s for *_, s in segs], reseeded, [s for *_, s in segs]):
            batch.run(seeds)
            assert batch.elapsed_ms > 0.0
            for k, ((P, p0, p1, R, _), block) in enumerate(zip(segs, seeds)):
                (o, r, c), reads = _single(vf, P, p0, p1, R, block)
                got = batch.edges(k)
                assert batch.nnz[k] == o.size, k
                assert all(np.array_equal(x, y) for x, y in zip(got, (o, r, c))), k
                assert got[2].dtype == np.int64
                assert batch.occupancy_reads[k] == reads, k
    finally:
        batch.close()


def test_batch_segment_matches_oracle(env, media):
    native, vf = env
    phases = _random(2, (12, 9, 11), 0.25)
    Q = O.build_surface_patches(phases, 0.004)
    P = media[1]
    p0, p1 = 10, Q.n_patches - 4
    batch = native.TraceBatch([(media[0].geometry(), 0, 40, 30, 1), (P.geometry(), p0, p1, 77, 1)],
                              ray_capacity=1 << 13)
    try:
        batch.run([[3], [123]])
        (o, r, c), reads = O.first_hit_counts_range(Q, 77, 123, p0, p1)
        got = batch.edges(1)
        assert all(np.array_equal(x, y) for x, y in zip(got, (o, r, c)))
        assert batch.occupancy_reads[1] == reads
    finally:
        batch.close()


def test_first_hit_counts_batch(env, media):
    _, vf = env
    work = [(media[4], 300, [21]), (media[0], 64, [22, 23]), (media[2], 9, [24])]
    got = vf.first_hit_counts_batch(work, ray_capacity=1 << 15)
    for (P, R, seeds), table in zip(work, got):
        ref = vf.first_hit_counts_pooled(P, R, seeds)
        assert all(np.array_equal(x, y) for x, y in zip(table, ref))


def test_edge_overflow_is_reported(env, media):
    native, _ = env
    P = media[0]
    batch = native.TraceBatch([(P.geometry(), 0, P.n_patches, 100, 1)], edge_capacity=1000)
    try:
        with pytest.raises(ValueError, match="edge capacity"):
            batch.run([[1]])
        with pytest.raises(ValueError, match="overflowed"):
            batch.edges(0)
    finally:
        batch.close()


def test_bad_segments_are_rejected(env, media):
    native, _ = env
    P = media[0]
    with pytest.raises(ValueError, match="out of bounds"):
        native.TraceBatch([(P.geometry(), 0, P.n_patches + 1, 10, 1)])
    with pytest.raises(ValueError, match="rays per patch"):
        native.TraceBatch([(P.geometry(), 0, 5, 10000, 2)])
    batch = native.TraceBatch([(P.geometry(), 0, 5, 10, 2)])
    try:
        with pytest.raises(ValueError, match="takes 2 seeds"):
            batch.run([[1]])
    finally:
        batch.close()


def test_all_archived_hash_cases_in_one_batch(env):
    native, vf = env
    hc = cases()["hash_cases"]
    patches = [vf.build_surface_patches(phases_x(h["microstructure"], h["axis"]), h["H_m"]) for h in hc]
    segments, seeds, expect = [], [], []
    for P, h in zip(patches, hc):
        R = h["rays_per_patch_per_block"]
        for which in "AB":
            segments.append((P.geometry(), 0, P.n_patches, R, len(h[f"seeds_{which}"])))
            seeds.append(h[f"seeds_{which}"])
            expect.append((P.n_patches, R * len(h[f"seeds_{which}"]), h[f"raw_count_sha256_{which}"]))
    batch = native.TraceBatch(segments)
    try:
        print("hash batch", batch.info, flush=True)
        batch.run(seeds)
        for k, (n, rays, sha) in enumerate(expect):
            o, r, c = batch.edges(k)
            assert vf.count_sha256(o, r, c, np.full(n, rays, np.int64)) == sha, k
    finally:
        batch.close()


def test_every_traversal_variant_in_batch_mode(env, media):
    native, vf = env
    segs = _segments(media)[:3]
    ref = [_single(vf, P, p0, p1, R, s)[0] for P, p0, p1, R, s in segs]
    variants = native.tracer_variants()
    default = next(v["variant"] for v in variants if v["current"])
    try:
        for v in variants:
            native.set_tracer_variant(v["variant"])
            batch = native.TraceBatch([(P.geometry(), p0, p1, R, len(s)) for P, p0, p1, R, s in segs],
                                      ray_capacity=1 << 15)
            try:
                batch.run([s for *_, s in segs])
                for k in range(len(segs)):
                    assert all(np.array_equal(x, y) for x, y in zip(batch.edges(k), ref[k])), (v, k)
            finally:
                batch.close()
    finally:
        native.set_tracer_variant(default)


def test_segment_edge_table_builds_the_same_operator(env, media):
    native, vf = env
    from porerad import exchange

    P = media[0]
    batch = native.TraceBatch([(media[1].geometry(), 0, 50, 20, 1), (P.geometry(), 0, P.n_patches, 40, 2)],
                              ray_capacity=1 << 15)
    try:
        batch.run([[9], [5, 6]])
        table = batch.edge_table(1)
        try:
            rays = np.full(P.n_patches, 80, np.int64)
            dev = exchange.ExchangeOperator.from_edges(table, rays, P.patch_area_m2)
        finally:
            table.close()
    finally:
        batch.close()
    o, r, c = vf.first_hit_counts_pooled(P, 40, [5, 6])
    host = exchange.ExchangeOperator(o, r, c, rays, P.patch_area_m2)
    assert np.array_equal(dev.scale, host.scale)
    x = np.random.default_rng(1).random(P.n_patches)
    assert np.array_equal(dev.rmatvec(x), host.rmatvec(x))
