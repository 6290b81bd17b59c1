"""CPU oracle of the porerad radiation hot path -- TEST INFRASTRUCTURE ONLY.

ctypes front end of ``libporerad_oracle.so`` (``porerad_oracle.c``), a plain
C restatement of the reference tracer (porerad/view_factors.py:70-323), plus a
NumPy restatement of the FP64 exchange operator and reflection series
(porerad/exchange.py:37-287) and of the prescribed-field channels
(porerad/coupled.py:45-194).  Only tests, the smoke check and the CPU-baseline
leg of the benchmark use it, as the checker; the product never imports it.

Pinning of each function (where the check lives):
  seed_state / uniforms ......... numpy.random.PCG64 / SeedSequence outputs
                                  (tests/test_oracle_rng.py)
  build_surface_patches ......... count hashes (patch ids enter the hash) and
                                  the closed-form patch counts of tiny grids
                                  (tests/test_oracle_tracer.py)
  generate_rays ................. count hashes; Lambertian moments
  trace_first_hit ............... count hashes; brute-force exact
                                  ray/plane intersection on tiny grids
  first_hit_counts / pool_counts  24 archived raw_count_sha256_A/B
  ExchangeOperator, reflection_series, prescribed_field_channels
                                  archived q_h/q_c/q_i of
                                  results/prescribed_field.csv to round-off;
                                  closed form of the empty box (parallel
                                  plates); isothermal zero net power
  mc_properties / mc_penetration  archived results/radiative_properties.csv
                                  (beta to round-off, omega and g within
                                  their standard errors) and
                                  results/penetration_response.csv (R, T, A
                                  within their standard errors); closed forms
                                  of the empty slab and of a solid wall
                                  (tests/test_oracle_mc.py)
"""

from __future__ import annotations

import ctypes
import hashlib
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libporerad_oracle.so")
_lib = None

HOT, COLD, INTERNAL = 0, 1, 2
SIGMA = 5.670374419e-8


def build() -> None:
    """Compile the oracle library in place (gcc, see Makefile)."""

    import subprocess

    subprocess.run(["make", "-s", "-C", _HERE, "-B"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.c_void_p
        i64 = ctypes.c_int64
        L.orc_seed_words.argtypes = [P, ctypes.c_int, P]
        L.orc_pcg64_advance.argtypes = [P, ctypes.c_uint64, ctypes.c_uint64]
        L.orc_pcg64_random.argtypes = [P, i64, P]
        L.orc_pcg64_raw.argtypes = [P, i64, P]
        L.orc_count_patches.argtypes = [P, i64, i64, i64, P]
        L.orc_count_patches.restype = i64
        L.orc_build_patches.argtypes = [P, i64, i64, i64] + [P] * 8
        L.orc_generate_rays.argtypes = [P, P, i64, i64, i64, i64, i64, P, P, P]
        L.orc_trace.argtypes = [P, i64, i64, i64, P, P, P, P, P, i64, P, P]
        L.orc_trace.restype = i64
        L.orc_first_hit_receivers.argtypes = [P, i64, i64, i64, P, P, P, P, P, i64, i64, i64, P, P, P, P]
        L.orc_first_hit_receivers.restype = i64
        L.orc_compact_edges.argtypes = [P, P, i64, i64, i64, P, P, P]
        L.orc_raw_matvec.argtypes = [P, P, P, i64, i64, P, P]
        L.orc_raw_rmatvec.argtypes = [P, P, P, i64, i64, P, P]
        L.orc_reflection_series_csr.argtypes = [P, P, P, i64, P, P, P, ctypes.c_double, i64, P, P, P, P]
        L.orc_reflection_series_csr.restype = i64
        d, u32, i32 = ctypes.c_double, ctypes.c_uint32, ctypes.c_int
        L.orc_mc_properties.argtypes = [P, i64, i64, i64, i64, u32, d, i32, i32, P, P, P, P, P, P]
        L.orc_mc_penetration.argtypes = [P, i64, i64, i64, i64, u32, d, i32, i32, P]
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


# ---------------------------------------------------------------- RNG

def seed_words(seed: int) -> np.ndarray:
    """uint32 little-endian words of a non-negative integer seed (SeedSequence entropy)."""

    seed = int(seed)
    if seed < 0:
        raise ValueError("expected non-negative integer")
    words = [0] if seed == 0 else []
    while seed > 0:
        words.append(seed & 0xFFFFFFFF)
        seed >>= 32
    return np.asarray(words, dtype=np.uint32)


def seed_state(seed: int) -> np.ndarray:
    """PCG64(seed) state as uint64 {state_hi, state_lo, inc_hi, inc_lo}."""

    words = seed_words(seed)
    out = np.zeros(4, dtype=np.uint64)
    lib().orc_seed_words(_p(words), int(words.size), _p(out))
    return out


def uniforms(seed: int, advance: int, count: int) -> np.ndarray:
    st = seed_state(seed)
    lib().orc_pcg64_advance(_p(st), ctypes.c_uint64(advance >> 64), ctypes.c_uint64(advance & (2**64 - 1)))
    out = np.empty(count, dtype=np.float64)
    lib().orc_pcg64_random(_p(st), count, _p(out))
    return out


def raw64(seed: int, count: int) -> np.ndarray:
    st = seed_state(seed)
    out = np.empty(count, dtype=np.uint64)
    lib().orc_pcg64_raw(_p(st), count, _p(out))
    return out


# ---------------------------------------------------------------- patches

@dataclass
class Patches:
    solid: np.ndarray
    patch_class: np.ndarray
    patch_cell: np.ndarray
    patch_face: np.ndarray
    patch_origin: np.ndarray
    patch_normal: np.ndarray
    patch_area_m2: np.ndarray
    cell_face_patch: np.ndarray
    hot_wall_patch: np.ndarray
    cold_wall_patch: np.ndarray
    H_m: float
    n_wall: int

    @property
    def n_patches(self) -> int:
        return int(self.patch_class.size)


def build_surface_patches(phases_x: np.ndarray, H_m: float) -> Patches:
    solid = np.ascontiguousarray((np.asarray(phases_x) == 1).astype(np.uint8))
    nx, ny, nz = map(int, solid.shape)
    n_wall = ctypes.c_int64(0)
    n = int(lib().orc_count_patches(_p(solid), nx, ny, nz, ctypes.byref(n_wall)))
    pc = np.empty(n, np.int8)
    cell = np.empty(n, np.int32)
    face = np.empty(n, np.int8)
    orig = np.empty((n, 3), np.float64)
    nrm = np.empty((n, 3), np.float64)
    cfp = np.empty((nx * ny * nz, 6), np.int32)
    hot = np.empty((ny, nz), np.int32)
    cold = np.empty((ny, nz), np.int32)
    lib().orc_build_patches(_p(solid), nx, ny, nz, _p(pc), _p(cell), _p(face), _p(orig), _p(nrm), _p(cfp),
                            _p(hot), _p(cold))
    area = (float(H_m) / nx) ** 2
    return Patches(solid, pc, cell, face, orig, nrm, np.full(n, area), cfp, hot, cold, float(H_m), int(n_wall.value))


# ---------------------------------------------------------------- tracer

def generate_rays(patches: Patches, p0: int, p1: int, rays_per_patch: int, seed: int):
    p0, p1, R = int(p0), int(p1), int(rays_per_patch)
    m = (p1 - p0) * R
    origins = np.empty((m, 3), np.float64)
    directions = np.empty((m, 3), np.float64)
    st = seed_state(seed)
    lib().orc_generate_rays(_p(patches.patch_origin), _p(patches.patch_normal), patches.solid.shape[1],
                            patches.solid.shape[2], p0, p1, R, _p(st), _p(origins), _p(directions))
    return origins, directions


def trace_rays(patches: Patches, origins: np.ndarray, directions: np.ndarray):
    """Receiver per ray (-1 = unresolved) and the total occupancy reads."""

    origins = np.ascontiguousarray(origins, dtype=np.float64)
    directions = np.ascontiguousarray(directions, dtype=np.float64)
    m = int(directions.shape[0])
    receiver = np.empty(m, np.int32)
    reads = ctypes.c_int64(0)
    nx, ny, nz = patches.solid.shape
    lib().orc_trace(_p(patches.solid), nx, ny, nz, _p(patches.hot_wall_patch), _p(patches.cold_wall_patch),
                    _p(patches.cell_face_patch), _p(origins), _p(directions), m, _p(receiver), ctypes.byref(reads))
    return receiver, int(reads.value)


def trace_first_hit(patches: Patches, origins: np.ndarray, directions: np.ndarray) -> np.ndarray:
    receiver, _ = trace_rays(patches, origins, directions)
    failed = int(np.count_nonzero(receiver < 0))
    if failed:
        raise RuntimeError(f"{failed} rays did not resolve a receiving patch")
    return receiver


def first_hit_counts_range(patches: Patches, rays_per_patch: int, seed: int, p0: int, p1: int):
    """Edges of patches [p0, p1) plus the occupancy-read total (exact algorithmic bytes)."""

    R = int(rays_per_patch)
    p0, p1 = int(p0), int(p1)
    work = np.empty((p1 - p0) * R, np.int32)
    n_unique = np.empty(p1 - p0, np.int64)
    reads = ctypes.c_int64(0)
    st = seed_state(seed)
    nx, ny, nz = patches.solid.shape
    failed = lib().orc_first_hit_receivers(
        _p(patches.solid), nx, ny, nz, _p(patches.hot_wall_patch), _p(patches.cold_wall_patch),
        _p(patches.cell_face_patch), _p(patches.patch_origin), _p(patches.patch_normal), p0, p1, R, _p(st),
        _p(work), _p(n_unique), ctypes.byref(reads))
    if failed:
        raise RuntimeError(f"{int(failed)} rays did not resolve a receiving patch")
    offsets = np.zeros(p1 - p0, np.int64)
    np.cumsum(n_unique[:-1], out=offsets[1:])
    nnz = int(n_unique.sum())
    origin = np.empty(nnz, np.int32)
    receiver = np.empty(nnz, np.int32)
    count = np.empty(nnz, np.int64)
    lib().orc_compact_edges(_p(work), _p(offsets), p0, p1, R, _p(origin), _p(receiver), _p(count))
    return (origin, receiver, count), int(reads.value)


def first_hit_counts(patches: Patches, rays_per_patch: int, seed: int):
    edges, _ = first_hit_counts_range(patches, rays_per_patch, seed, 0, patches.n_patches)
    return edges


def pool_counts(tables, n_patches: int):
    n = int(n_patches)
    pair = np.concatenate([o.astype(np.int64) * n + r.astype(np.int64) for o, r, _ in tables])
    count = np.concatenate([np.asarray(c, dtype=np.int64) for _, _, c in tables])
    order = np.argsort(pair, kind="stable")
    pair, count = pair[order], count[order]
    start = np.r_[0, np.flatnonzero(pair[1:] != pair[:-1]) + 1]
    unique = pair[start]
    return (unique // n).astype(np.int32), (unique % n).astype(np.int32), np.add.reduceat(count, start)


def count_sha256(origin, receiver, count, rays_per_origin) -> str:
    digest = hashlib.sha256()
    for name, array in (
        ("origin_patch_id", np.asarray(origin, dtype=np.int32)),
        ("receiver_patch_id", np.asarray(receiver, dtype=np.int32)),
        ("count", np.asarray(count, dtype=np.int64)),
        ("rays_per_origin", np.asarray(rays_per_origin, dtype=np.int64)),
    ):
        digest.update(name.encode("ascii"))
        digest.update(array.dtype.str.encode("ascii"))
        digest.update(np.asarray(array.shape, dtype=np.int64).tobytes())
        digest.update(np.ascontiguousarray(array).tobytes())
    return digest.hexdigest()


def orient_along_x(array: np.ndarray, axis: str) -> np.ndarray:
    return np.ascontiguousarray(np.moveaxis(np.asarray(array), {"x": 0, "y": 1, "z": 2}[axis], 0))


def mc_properties(solid_x: np.ndarray, n_rays: int, seed: int, epsilon: float, max_bounces: int = 1000,
                  restart: str = "article") -> dict:
    """Per-ray first path lengths and reflection statistics of trace_properties's
    ray kernel (porerad/properties.py:223-287) for one seed."""

    solid = np.ascontiguousarray(solid_x, dtype=np.uint8)
    n = int(n_rays)
    ev, ce, m1, m2 = (np.empty(n) for _ in range(4))
    inter, scat = np.empty(n, np.int64), np.empty(n, np.int64)
    lib().orc_mc_properties(_p(solid), *solid.shape, n, int(seed) & 0xFFFFFFFF, float(epsilon), int(max_bounces),
                            1 if restart == "hit_point" else 0, _p(ev), _p(ce), _p(inter), _p(scat), _p(m1), _p(m2))
    return {"event": ev, "censored": ce, "interactions": inter, "scattered": scat, "mu": m1, "mu2": m2}


def mc_penetration(solid_x: np.ndarray, n_rays: int, seed: int, epsilon: float, max_bounces: int = 1000,
                   restart: str = "article") -> np.ndarray:
    """Per-ray fate of trace_penetration's ray kernel (porerad/penetration.py:200-272):
    0 reflected, 1 transmitted, 2 absorbed, 3 stopped."""

    solid = np.ascontiguousarray(solid_x, dtype=np.uint8)
    fate = np.empty(int(n_rays), np.uint8)
    lib().orc_mc_penetration(_p(solid), *solid.shape, int(n_rays), int(seed) & 0xFFFFFFFF, float(epsilon),
                             int(max_bounces), 1 if restart == "hit_point" else 0, _p(fate))
    return fate
