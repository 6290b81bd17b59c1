"""Host-side logic and the C ABI library, without GPU compute.

* libporerad_hip.so loads and exports every entry point declared in
  include/porerad_hip.h, and the ctypes table declares each of them.
* prh_pcg64_seed (host code of the library) equals numpy's PCG64 seeding.
* count_sha256 / pool_counts of the drop-in package equal the oracle's and
  hashlib's digests.
"""

import ctypes
import os
import re

import numpy as np
import pytest

from oracle import oracle as O

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "porerad_hip.h")


def _declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(prh_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from porerad import _native

    assert os.path.exists(_native.LIB_PATH), "build the library first (__graft_entry__.build())"
    raw = ctypes.CDLL(_native.LIB_PATH)
    declared = _declared()
    assert len(declared) >= 18
    for name in declared:
        assert hasattr(raw, name), name
        assert name in _native.SIGNATURES, name
    _native.load()  # every declared signature binds


def test_library_seeding_matches_numpy():
    from porerad import _native

    for seed in [0, 5, 2**40 + 1, 4734788061198622954, 2**70 + 3]:
        st = np.random.PCG64(seed).state["state"]
        ref = [st["state"] >> 64, st["state"] & (2**64 - 1), st["inc"] >> 64, st["inc"] & (2**64 - 1)]
        assert [int(v) for v in _native.pcg64_state(seed)] == ref


def test_count_sha256_matches_oracle_and_hashlib():
    import hashlib

    from porerad import view_factors as vf

    o = np.array([0, 0, 1, 2], np.int32)
    r = np.array([1, 2, 0, 0], np.int32)
    c = np.array([3, 1, 4, 4], np.int64)
    rays = np.full(3, 4, np.int64)
    h = hashlib.sha256()
    for name, a in (("origin_patch_id", o), ("receiver_patch_id", r), ("count", c), ("rays_per_origin", rays)):
        h.update(name.encode())
        h.update(a.dtype.str.encode())
        h.update(np.asarray(a.shape, np.int64).tobytes())
        h.update(a.tobytes())
    assert vf.count_sha256(o, r, c, rays) == h.hexdigest() == O.count_sha256(o, r, c, rays)


def test_pool_counts_matches_oracle():
    from porerad import view_factors as vf

    rng = np.random.default_rng(0)
    tables = []
    for _ in range(3):
        o = np.sort(rng.integers(0, 50, 400)).astype(np.int32)
        r = rng.integers(0, 50, 400).astype(np.int32)
        c = rng.integers(1, 5, 400).astype(np.int64)
        tables.append((o, r, c))
    a = vf.pool_counts(tables, 50)
    b = O.pool_counts(tables, 50)
    for x, y in zip(a, b):
        assert x.dtype == y.dtype and np.array_equal(x, y)


def test_native_errors_map_to_reference_exceptions():
    from porerad import _native

    with pytest.raises(ValueError):
        _native.seed_words(-1)


def _fibres_brute_force(N, solid_fraction, radius, seed):
    """Every voxel centre against every fibre (the definition of random_fibers)."""

    rng = np.random.default_rng(seed)
    c = np.arange(N) + 0.5
    P = np.stack(np.meshgrid(c, c, c, indexing="ij"), axis=-1).reshape(-1, 3)
    solid = np.zeros(len(P), bool)
    while solid.mean() < solid_fraction:
        centre = rng.uniform(0.0, N, size=3)
        axis = rng.normal(size=3)
        axis /= np.linalg.norm(axis)
        v = P - centre
        v -= N * np.round(v / N)
        t = np.clip(v @ axis, -N / 2.0, N / 2.0)
        solid |= np.sum((v - t[:, None] * axis) ** 2, axis=1) <= radius * radius
    return solid.reshape(N, N, N).astype(np.uint8)


@pytest.mark.parametrize("N,radius,seed", [(24, 3.0, 0), (40, 2.5, 4), (64, 3.0, 0), (33, 5.0, 9)])
def test_random_fibres_tube_windows_equal_brute_force(N, radius, seed):
    from porerad import synthetic

    assert np.array_equal(synthetic.random_fibers(N, 0.10, radius, seed), _fibres_brute_force(N, 0.10, radius, seed))


def test_case_driver_helpers(tmp_path):
    from porerad import cases as C

    assert C._seeds("5;7;11") == [5, 7, 11] and C._seeds([3, 4]) == [3, 4]
    assert C._conditions("400/0.3;700/0.9") == [(400.0, 0.3), (700.0, 0.9)]
    out = tmp_path / "rows.csv"
    first = [{"case_id": "a", "q_h_W_m2": 1.0}, {"case_id": "b", "q_h_W_m2": 2.0}]
    C.write_rows(out, first, ["case_id", "q_h_W_m2"])
    C.write_rows(out, [{"case_id": "a", "q_h_W_m2": 3.0}], ["case_id", "q_h_W_m2"])  # replaces case a only
    import csv

    with open(out, newline="") as handle:
        rows = list(csv.DictReader(handle))
    assert [(r["case_id"], r["q_h_W_m2"]) for r in rows] == [("b", "2.0"), ("a", "3.0")]
    with pytest.raises(ValueError, match="differs from the one listed"):
        C._check_phases(np.zeros((2, 2, 2), np.uint8), "0" * 64)


@pytest.mark.parametrize("bits", [1, 2, 4])
def test_hilbert_keys_walk_the_grid_one_cell_at_a_time(bits):
    """The sweep order's Hilbert keys number a full 2^b cube once each, and
    consecutive keys are face neighbours (what makes the curve local)."""

    from porerad.view_factors import _hilbert_keys

    m = 1 << bits
    grid = np.stack(np.meshgrid(*(np.arange(m),) * 3, indexing="ij"), -1).reshape(-1, 3)
    key = _hilbert_keys(grid, bits)
    assert np.array_equal(np.sort(key), np.arange(m ** 3))
    walk = grid[np.argsort(key)]
    if m > 1:
        assert np.abs(np.diff(walk, axis=0)).sum(axis=1).max() == 1


def test_sweep_orders_are_permutations():
    from porerad import view_factors as vf

    rng = np.random.default_rng(3)
    phases = (rng.random((10, 8, 9)) < 0.2).astype(np.uint8)
    Q = O.build_surface_patches(phases, 0.01)
    P = vf.SurfacePatches(Q.solid, Q.patch_class, Q.patch_cell, Q.patch_face, Q.patch_origin, Q.patch_normal,
                          Q.patch_area_m2, Q.cell_face_patch, Q.hot_wall_patch, Q.cold_wall_patch, Q.H_m)
    for curve in ("morton", "hilbert"):
        order = P.sweep_order(curve)
        assert order.dtype == np.int32 and np.array_equal(np.sort(order), np.arange(P.n_patches))
    with pytest.raises(ValueError):
        P.sweep_order("peano")
