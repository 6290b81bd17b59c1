"""Parity of the HIP tracer with the CPU oracle and the reference's archived hashes.

All calls go through the C ABI (libporerad_hip.so via porerad._native).
Integer outputs must be bit-identical: patch tables, receivers, counts.
"""

from __future__ import annotations

import numpy as np
import pytest

from golden_data import cases, phases_x
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def vf():
    from porerad import _native, view_factors

    _native.lib()
    return view_factors


_cache = {}


def _both(vf, key, phases, H=0.01):
    if key not in _cache:
        _cache.clear()
        _cache[key] = (vf.build_surface_patches(phases, H), O.build_surface_patches(phases, H))
    return _cache[key]


def _fiber_c1():
    from porerad import synthetic

    return synthetic.random_fibers(64, 0.10, 3.0, 0)


def _small_random(seed=0, shape=(12, 10, 11), density=0.2):
    rng = np.random.default_rng(seed)
    return (rng.random(shape) < density).astype(np.uint8)


@pytest.mark.parametrize("name,axis", [("OF_canonical_phi0.10_N128", "x"), ("FI_canonical_phi0.10_N128", "z"),
                                       ("synthetic_fiber_N64", "x"), ("small_random", "x")])
def test_patch_table_matches_oracle(vf, name, axis):
    if name == "synthetic_fiber_N64":
        phases = _fiber_c1()
    elif name == "small_random":
        phases = _small_random()
    else:
        phases = phases_x(name, axis)
    P, Q = _both(vf, (name, axis), phases)
    assert P.n_patches == Q.n_patches
    for field in ("patch_class", "patch_cell", "patch_face", "patch_origin", "patch_normal", "cell_face_patch",
                  "hot_wall_patch", "cold_wall_patch"):
        a, b = getattr(P, field), getattr(Q, field)
        assert a.dtype == b.dtype or field == "patch_normal", field
        assert np.array_equal(a, b), field


def test_generate_rays_matches_oracle(vf):
    phases = phases_x("OF_canonical_phi0.10_N128", "x")
    P, Q = _both(vf, ("OF_canonical_phi0.10_N128", "x"), phases)
    n = P.n_patches
    for p0, p1 in ((0, 3000), (n - 3000, n)):
        o_gpu, d_gpu = vf.generate_rays(P, p0, p1, 64, 4734788061198622954)
        o_ref, d_ref = O.generate_rays(Q, p0, p1, 64, 4734788061198622954)
        assert np.array_equal(o_gpu, o_ref)
        # directions bit for bit: the device sin/cos restates the C library's
        assert np.array_equal(d_gpu.view(np.int64), d_ref.view(np.int64))


def test_trace_given_rays_matches_oracle(vf):
    phases = phases_x("FA_canonical_phi0.10_N128", "y")
    P, Q = _both(vf, ("FA_canonical_phi0.10_N128", "y"), phases)
    o, d = O.generate_rays(Q, 0, P.n_patches, 4, 17)
    rec_ref, _ = O.trace_rays(Q, o, d)
    assert np.all(rec_ref >= 0)
    rec = vf.trace_first_hit(P, o, d)
    assert np.array_equal(rec, rec_ref)


def _device_receivers(P, o, d):
    """prh_trace_rays' receivers with -1 for unresolved rays (no exception)."""

    from porerad import _native

    o = np.ascontiguousarray(o, dtype=np.float64)
    d = np.ascontiguousarray(d, dtype=np.float64)
    receiver = np.empty(d.shape[0], dtype=np.int32)
    failed = np.zeros(1, dtype=np.int64)
    status = _native.lib().prh_trace_rays(P.geometry().handle, _native.ptr(o), _native.ptr(d), int(d.shape[0]),
                                          _native.ptr(receiver), _native.ptr(failed))
    assert status in (_native.PRH_OK, _native.PRH_ERR_UNRESOLVED)
    assert int(failed[0]) == int(np.count_nonzero(receiver < 0))
    return receiver


def test_trace_origins_outside_the_box_match_oracle(vf):
    # The reference reads j % ny, k % nz of unbounded indices and sends any
    # ray whose x index starts below 0 (above nx - 1) to the hot (cold) wall.
    phases = phases_x("TL_canonical_phi0.10_N128", "x")
    P, Q = _both(vf, ("TL_canonical_phi0.10_N128", "x"), phases)
    nx, ny, nz = phases.shape
    o, d = O.generate_rays(Q, 0, P.n_patches, 2, 91)
    rng = np.random.default_rng(4)
    shifted = o.copy()
    shifted[:, 1] += ny * rng.integers(-40, 41, size=o.shape[0])
    shifted[:, 2] += nz * rng.integers(-40, 41, size=o.shape[0])
    far = rng.random(o.shape[0]) < 0.05
    shifted[far, 0] = rng.choice([-1000.5, -1.5, -1.0, nx + 0.25, nx + 700.0], size=int(far.sum()))
    ref, _ = O.trace_rays(Q, shifted, d)
    got = _device_receivers(P, shifted, d)
    assert np.array_equal(got, ref)
    assert np.count_nonzero(ref >= 0) > 0.9 * ref.size


def test_trace_rejects_occupancy_only_geometry_and_bad_rays(vf):
    from porerad import _native

    phases = _small_random()
    G = _native.Geometry.occupancy_only(phases)
    o = np.full((1, 3), 0.5)
    d = np.array([[1.0, 0.0, 0.0]])
    receiver = np.empty(1, np.int32)
    status = _native.lib().prh_trace_rays(G.handle, _native.ptr(o), _native.ptr(d), 1, _native.ptr(receiver), None)
    assert status == _native.PRH_ERR_ARG
    assert "patch tables" in _native.last_error()
    P = vf.build_surface_patches(phases, 0.01)
    for bad_o, bad_d in ((np.array([[np.nan, 1.0, 1.0]]), d), (np.array([[0.5, 3e9, 1.0]]), d),
                         (o, np.array([[1.0, np.inf, 0.0]]))):
        with pytest.raises(ValueError):
            vf.trace_first_hit(P, bad_o, bad_d)


def _equal_tables(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a, b))


def test_counts_tiled_under_a_small_workspace_equal_one_pass(vf):
    from porerad import _native

    phases = _fiber_c1()
    P, Q = _both(vf, ("synthetic_fiber_N64", "x"), phases)
    one = vf.first_hit_counts_pooled(P, 200, [1, 2])
    _, limit = _native.tracer_workspace()
    try:
        # 8 MiB holds about 180 patches of 400 rays: the range is traced in ~180 passes
        _native.set_tracer_workspace_limit(8 << 20)
        tiled = vf.first_hit_counts_pooled(P, 200, [1, 2])
        held, _ = _native.tracer_workspace()
        assert held <= 8 << 20
    finally:
        _native.set_tracer_workspace_limit(limit)
    assert _equal_tables(tiled, one)
    ref = O.pool_counts([O.first_hit_counts(Q, 200, s) for s in (1, 2)], Q.n_patches)
    assert _equal_tables(one, ref)


@pytest.mark.parametrize("R,seeds", [(6000, [5, 6, 7]), (40000, [9]), (20000, [3, 4])])
def test_counts_beyond_one_pass_per_patch_match_oracle(vf, R, seeds):
    # more than 16384 rays per patch: seed groups, or parts of one seed's rays, pooled on the device
    phases = _small_random(seed=3, shape=(10, 9, 8), density=0.25)
    P, Q = _both(vf, ("small_random_3", "x"), phases)
    got = vf.first_hit_counts_pooled(P, R, seeds)
    ref = O.pool_counts([O.first_hit_counts(Q, R, s) for s in seeds], Q.n_patches)
    assert _equal_tables(got, ref)
    assert int(got[2].sum()) == P.n_patches * R * len(seeds)


def _hash_params():
    out = []
    for h in cases()["hash_cases"]:
        for s in "AB":
            out.append(pytest.param(h, s, id=f"{h['structure']}-{h['axis']}-{s}"))
    return out


@pytest.mark.parametrize("case,which", _hash_params())
def test_counts_reproduce_archived_hash(vf, case, which):
    phases = phases_x(case["microstructure"], case["axis"])
    key = (case["microstructure"], case["axis"])
    if key not in _cache:
        _cache.clear()
        _cache[key] = (vf.build_surface_patches(phases, case["H_m"]), None)
    P = _cache[key][0]
    R = case["rays_per_patch_per_block"]
    seeds = case[f"seeds_{which}"]
    o, r, c = vf.first_hit_counts_pooled(P, R, seeds)
    rays = np.full(P.n_patches, R * len(seeds), dtype=np.int64)
    assert vf.count_sha256(o, r, c, rays) == case[f"raw_count_sha256_{which}"]


def test_c1_fibre_counts_and_reads_bit_exact(vf):
    from porerad import _native

    phases = _fiber_c1()
    P, Q = _both(vf, ("synthetic_fiber_N64", "x"), phases)
    edges = vf.first_hit_edges(P, 200, [1])
    got = edges.to_host()
    reads = edges.occupancy_reads
    edges.close()
    (o, r, c), reads_ref = O.first_hit_counts_range(Q, 200, 1, 0, Q.n_patches)
    assert np.array_equal(got[0], o) and np.array_equal(got[1], r) and np.array_equal(got[2], c)
    assert got[2].dtype == np.int64
    assert reads == reads_ref
    assert np.all(_native.last_timings()[:2] > 0)


def test_single_seed_api_equals_pooled_singles(vf):
    phases = _small_random(3, (14, 12, 12), 0.15)
    P, Q = _both(vf, ("small3", "x"), phases)
    seeds = [5, 99, 2**63 + 7]
    tables = [vf.first_hit_counts(P, 37, s) for s in seeds]
    pooled = vf.first_hit_counts_pooled(P, 37, seeds)
    ref = O.pool_counts([O.first_hit_counts(Q, 37, s) for s in seeds], Q.n_patches)
    for a, b, c in zip(vf.pool_counts(tables, P.n_patches), pooled, ref):
        assert np.array_equal(a, c) and np.array_equal(b, c)


@pytest.mark.parametrize("R", [1, 7, 64, 1000, 2000, 4096])
def test_partial_ranges_and_ray_counts(vf, R):
    phases = _small_random(4, (16, 9, 13), 0.2)
    P, Q = _both(vf, ("small4", "x"), phases)
    n = P.n_patches
    for p0, p1 in ((0, n), (n // 3, n // 3 + 17), (n - 1, n), (5, 5)):
        got = vf.first_hit_counts_pooled(P, R, [123], patch_range=(p0, p1))
        (o, r, c), _ = O.first_hit_counts_range(Q, R, 123, p0, p1)
        assert all(np.array_equal(x, y) for x, y in zip(got, (o, r, c))), (p0, p1)


@pytest.mark.parametrize("R,seeds", [(1000, [11, 12]), (500, [1, 2, 3, 4])])
def test_pooled_seeds_up_to_2048_rays_per_patch_match_oracle(vf, R, seeds):
    # 1025..2048 rays per patch and call: one wavefront per patch, 32 keys per lane
    phases = _small_random(6, (12, 10, 11), 0.2)
    P, Q = _both(vf, ("small6", "x"), phases)
    got = vf.first_hit_counts_pooled(P, R, seeds)
    ref = O.pool_counts([O.first_hit_counts(Q, R, s) for s in seeds], Q.n_patches)
    assert _equal_tables(got, ref)
    assert int(got[2].sum()) == P.n_patches * R * len(seeds)


def test_counts_are_deterministic(vf):
    phases = _small_random(5, (20, 16, 16), 0.12)
    P, _ = _both(vf, ("small5", "x"), phases)
    a = vf.first_hit_counts_pooled(P, 300, [8, 9])
    b = vf.first_hit_counts_pooled(P, 300, [8, 9])
    assert all(np.array_equal(x, y) for x, y in zip(a, b))


def test_unresolved_rays_raise(vf):
    phases = np.zeros((4, 4, 4), np.uint8)
    phases[1, 1, 1] = 1
    P = vf.build_surface_patches(phases, 0.004)
    o = np.array([[0.5, 0.5, 0.5], [1.5, 1.5, 1.5], [0.5, 0.5, 0.5]])
    d = np.array([[0.0, 1.0, 0.0], [1.0, 0.0, 0.0], [1.0, 0.0, 0.0]])
    with pytest.raises(RuntimeError, match="2 rays did not resolve a receiving patch"):
        vf.trace_first_hit(P, o, d)


def test_empty_box_counts(vf):
    phases = np.zeros((8, 6, 6), np.uint8)
    P, Q = _both(vf, ("empty", "x"), phases, 0.008)
    got = vf.first_hit_counts(P, 50, 3)
    ref = O.first_hit_counts(Q, 50, 3)
    assert all(np.array_equal(x, y) for x, y in zip(got, ref))
    assert np.all(P.patch_class[got[0]] != P.patch_class[got[1]])


@pytest.mark.parametrize("shape", [(24, 20, 20), (20, 16, 32)], ids=["lateral-20x20", "lateral-pow2"])
def test_every_traversal_variant_gives_identical_counts(vf, shape):
    """Counts and the occupancy-read total of every variant against the oracle,
    on lateral sizes that are and are not powers of two (separate kernels)."""

    from porerad import _native

    phases = _small_random(6, shape, 0.12)
    P, Q = _both(vf, ("small6", shape), phases)
    ref, reads_ref = O.first_hit_counts_range(Q, 300, 77, 0, Q.n_patches)
    variants = _native.tracer_variants()
    assert len(variants) >= 2
    default = next(v["variant"] for v in variants if v["current"])
    try:
        for v in variants:
            _native.set_tracer_variant(v["variant"])
            edges = vf.first_hit_edges(P, 300, [77])
            got, reads = edges.to_host(), edges.occupancy_reads
            edges.close()
            assert all(np.array_equal(a, b) for a, b in zip(got, ref)), v
            assert reads == reads_ref, v
    finally:
        _native.set_tracer_variant(default)
