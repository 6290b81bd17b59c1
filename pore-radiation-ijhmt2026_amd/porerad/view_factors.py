"""Surface patches, Lambertian ray sets and first-hit exchange counts on MI355X.

Drop-in for the reference module porerad/view_factors.py: same names,
signatures, dtypes, orderings and exceptions.  The patch table is built on
the GPU (``prh_geometry_create``) and downloaded into the ``SurfacePatches``
arrays; ray generation, the FP64 voxel traversal and the per-patch integer
counts run in hand-written HIP kernels (``prh_generate_rays``,
``prh_trace_rays``, ``prh_trace_counts``).  The counts are bit-identical to
the reference CPU implementation for the same seeds.

Conventions kept from the reference: the transfer axis is x (hot wall at
x = 0, cold wall at x = N_x), y and z are periodic, positions are in cell
widths, patch classes are HOT = 0, COLD = 1, INTERNAL = 2, and faces 0..5 are
+x, -x, +y, -y, +z, -z.
"""

from __future__ import annotations

import hashlib
from dataclasses import dataclass, field

import numpy as np

from . import _native

HOT, COLD, INTERNAL = 0, 1, 2
"""Patch classes: hot wall, cold wall, solid-pore interface (view_factors.py:25)."""

_AXIS_INDEX = {"x": 0, "y": 1, "z": 2}
_FACE_NORMALS = np.asarray(
    ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)), dtype=np.float64
)


def orient_along_x(array: np.ndarray, axis: str) -> np.ndarray:
    """The array with the transfer axis moved to index 0, C-contiguous (view_factors.py:37-40)."""

    return np.ascontiguousarray(np.moveaxis(np.asarray(array), _AXIS_INDEX[axis], 0))


def _hilbert_keys(ijk: np.ndarray, bits: int) -> np.ndarray:
    """Index along the 3-D Hilbert curve of order `bits` of each integer point
    (n x 3, coordinates below 2**bits): Skilling's transform from axes to the
    transposed index, then the bits interleaved."""

    X = np.array(ijk, dtype=np.int64, copy=True)
    top = 1 << (bits - 1)
    q = top
    while q > 1:
        low = q - 1
        for i in range(3):
            up = (X[:, i] & q) != 0
            X[up, 0] ^= low
            down = ~up
            t = (X[down, 0] ^ X[down, i]) & low
            X[down, 0] ^= t
            X[down, i] ^= t
        q >>= 1
    X[:, 1] ^= X[:, 0]
    X[:, 2] ^= X[:, 1]
    t = np.zeros(X.shape[0], np.int64)
    q = top
    while q > 1:
        t[(X[:, 2] & q) != 0] ^= q - 1
        q >>= 1
    X ^= t[:, None]
    key = np.zeros(X.shape[0], np.int64)
    for b in range(bits - 1, -1, -1):
        for i in range(3):
            key = (key << 1) | ((X[:, i] >> b) & 1)
    return key


@dataclass
class SurfacePatches:
    """Patch table of one x-oriented microstructure (view_factors.py:43-67).

    The arrays are host copies of the device table; ``device`` holds the
    resident GPU geometry used by the tracer.
    """

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
    device: _native.Geometry | None = field(default=None, repr=False, compare=False)
    _owners: tuple | None = field(default=None, init=False, repr=False, compare=False)

    @property
    def n_patches(self) -> int:
        return int(self.patch_class.size)

    def sweep_order(self, curve: str = "morton") -> np.ndarray:
        """The patches along a space-filling curve of their origins on the
        half-voxel grid ("morton": Z-order; "hilbert"): an order in which
        consecutive patches are close in space, for the batched engine's row
        sweep (see ``ExchangeOperator.set_sweep_order``)."""

        ijk = np.floor(np.asarray(self.patch_origin) * 2.0).astype(np.int64)
        ijk -= ijk.min(axis=0)
        if curve == "hilbert":
            bits = max(1, int(ijk.max()).bit_length())
            return np.argsort(_hilbert_keys(ijk, bits), kind="stable").astype(np.int32)
        if curve != "morton":
            raise ValueError(f"unknown curve {curve!r}")
        key = np.zeros(self.n_patches, np.int64)
        for b in range(20):
            key |= (((ijk[:, 0] >> b) & 1) << (3 * b + 2)) | (((ijk[:, 1] >> b) & 1) << (3 * b + 1)) | \
                   (((ijk[:, 2] >> b) & 1) << (3 * b))
        return np.argsort(key, kind="stable").astype(np.int32)

    def internal_owners(self) -> tuple[np.ndarray, np.ndarray]:
        """(mask of the internal patches, flat index of the solid cell owning each), built once."""

        if self._owners is None:
            inner = self.patch_class == INTERNAL
            self._owners = (inner, self.patch_cell[inner].astype(np.intp))
        return self._owners

    def geometry(self) -> _native.Geometry:
        """The device patch table, uploaded on first use."""

        if self.device is None:
            self.device = _native.Geometry(self.solid, self.H_m)
        return self.device


def build_surface_patches(phases_x: np.ndarray, H_m: float) -> SurfacePatches:
    """Wall and interface patches of an x-oriented phase array, 1 = solid (view_factors.py:70-132).

    Wall patches come first in (j, k) order, hot before cold, where the
    boundary cell is pore; internal patches follow in (i, j, k, face) order.
    All patches have the area (H/N_x)^2.
    """

    solid = np.ascontiguousarray((np.asarray(phases_x) == 1).astype(np.uint8))
    if solid.ndim != 3:
        raise ValueError("expected a 3-D phase array")
    nx, ny, nz = map(int, solid.shape)
    geo = _native.Geometry(solid, float(H_m))
    patch_class, patch_cell, patch_face, patch_origin = geo.patch_table()
    n = geo.n_patches

    normal = np.zeros((n, 3), dtype=np.float64)
    wall = patch_class != INTERNAL
    normal[wall, 0] = np.where(patch_class[wall] == HOT, 1.0, -1.0)
    normal[~wall] = _FACE_NORMALS[patch_face[~wall]]

    cell_face_patch = np.full((nx * ny * nz, 6), -1, dtype=np.int32)
    ids = np.arange(n, dtype=np.int32)
    cell_face_patch[patch_cell[~wall], patch_face[~wall]] = ids[~wall]
    hot_wall_patch = np.full((ny, nz), -1, dtype=np.int32)
    cold_wall_patch = np.full((ny, nz), -1, dtype=np.int32)
    j = np.floor(patch_origin[wall, 1]).astype(np.int64)
    k = np.floor(patch_origin[wall, 2]).astype(np.int64)
    is_hot = patch_class[wall] == HOT
    hot_wall_patch[j[is_hot], k[is_hot]] = ids[wall][is_hot]
    cold_wall_patch[j[~is_hot], k[~is_hot]] = ids[wall][~is_hot]

    return SurfacePatches(
        solid=solid,
        patch_class=patch_class,
        patch_cell=patch_cell,
        patch_face=patch_face,
        patch_origin=patch_origin,
        patch_normal=normal,
        patch_area_m2=np.full(n, (float(H_m) / nx) ** 2, dtype=np.float64),
        cell_face_patch=cell_face_patch,
        hot_wall_patch=hot_wall_patch,
        cold_wall_patch=cold_wall_patch,
        H_m=float(H_m),
        device=geo,
    )


def generate_rays(
    patches: SurfacePatches, patch_start: int, patch_end: int, rays_per_patch: int, seed: int
) -> tuple[np.ndarray, np.ndarray]:
    """Origins and directions of the rays of patches [patch_start, patch_end) (view_factors.py:135-182).

    Ray r of patch p uses draws 4 (p R + r) .. 4 (p R + r) + 3 of the
    PCG64(seed) stream, so any partition reproduces the same rays.
    """

    p0, p1, R = int(patch_start), int(patch_end), int(rays_per_patch)
    if not 0 <= p0 <= p1 <= patches.n_patches:
        raise ValueError("patch range out of bounds")
    m = (p1 - p0) * R
    origins = np.empty((m, 3), dtype=np.float64)
    directions = np.empty((m, 3), dtype=np.float64)
    if m:
        state = _native.pcg64_state(seed)
        _native.check(_native.lib().prh_generate_rays(patches.geometry().handle, p0, p1, R, _native.ptr(state),
                                                      _native.ptr(origins), _native.ptr(directions)),
                      "prh_generate_rays")
    return origins, directions


def trace_first_hit(patches: SurfacePatches, origins: np.ndarray, directions: np.ndarray) -> np.ndarray:
    """Index of the first patch hit by each ray (view_factors.py:261-270).

    Raises RuntimeError("{k} rays did not resolve a receiving patch") when a
    ray starts in a solid cell, has dx == 0, hits an unexposed face, or
    exceeds the step cap.
    """

    o = np.ascontiguousarray(origins, dtype=np.float64).reshape(-1, 3)
    d = np.ascontiguousarray(directions, dtype=np.float64).reshape(-1, 3)
    if o.shape != d.shape:
        raise ValueError("origins and directions must have the same shape")
    receiver = np.empty(d.shape[0], dtype=np.int32)
    failed = np.zeros(1, dtype=np.int64)
    status = _native.lib().prh_trace_rays(patches.geometry().handle, _native.ptr(o), _native.ptr(d), int(d.shape[0]),
                                          _native.ptr(receiver), _native.ptr(failed))
    if status == _native.PRH_ERR_UNRESOLVED:
        raise RuntimeError(f"{int(failed[0])} rays did not resolve a receiving patch")
    _native.check(status, "prh_trace_rays")
    return receiver


def first_hit_counts(
    patches: SurfacePatches, rays_per_patch: int, seed: int, *, patch_chunk: int = 16384
) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Integer counts (origin, receiver, n) of one seeded ray block, sorted by edge (view_factors.py:273-287).

    The patch range is traced in fused device passes; the library cuts it
    into as many passes as its bounded workspace needs
    (``_native.set_tracer_workspace_limit``), which is what ``patch_chunk``
    does for the reference's host arrays.  ``patch_chunk`` is accepted and
    must be positive; like the reference's, it never changes the result (the
    ray stream is partition invariant).
    """

    if int(patch_chunk) <= 0:
        raise ValueError("patch_chunk must be positive")

    return first_hit_counts_pooled(patches, rays_per_patch, [seed])


def first_hit_counts_pooled(
    patches: SurfacePatches, rays_per_patch: int, seeds, *, patch_range: tuple[int, int] | None = None
) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Counts of several seeded ray blocks pooled in one fused device pass.

    Equal to ``pool_counts([first_hit_counts(patches, R, s) for s in seeds], n)``.
    """

    edges = first_hit_edges(patches, rays_per_patch, seeds, patch_range=patch_range)
    try:
        return edges.to_host()
    finally:
        edges.close()


def first_hit_edges(patches: SurfacePatches, rays_per_patch: int, seeds, *,
                    patch_range: tuple[int, int] | None = None) -> _native.Edges:
    """Device-resident pooled edge table (kept in HBM for the operator build)."""

    seeds = [int(s) for s in seeds]
    if not seeds:
        raise ValueError("need at least one seed")
    p0, p1 = patch_range if patch_range is not None else (0, patches.n_patches)
    return _native.trace_counts(patches.geometry(), p0, p1, int(rays_per_patch), seeds)


def first_hit_counts_batch(cases, *, ray_capacity: int = 0, edge_capacity: int = 0):
    """Counts of many geometries in one graph-captured device pass.

    ``cases`` lists ``(patches, rays_per_patch, seeds)``; the result lists one
    ``(origin, receiver, count)`` per case, identical to
    ``first_hit_counts_pooled(patches, rays_per_patch, seeds)``.  This is the
    per-case loop of scripts/run_prescribed_field.py:81-85 in one replay:
    uneven geometries share persistent launches, in chunks of at most
    ``ray_capacity`` ray records (0: up to 2^28).  For repeated passes over the
    same layout keep a :class:`porerad._native.TraceBatch` and call ``run``.
    """

    segments, seeds = [], []
    for patches, rays_per_patch, block in cases:
        block = [int(s) for s in block]
        if not block:
            raise ValueError("need at least one seed per case")
        segments.append((patches.geometry(), 0, patches.n_patches, int(rays_per_patch), len(block)))
        seeds.append(block)
    batch = _native.TraceBatch(segments, ray_capacity=ray_capacity, edge_capacity=edge_capacity)
    try:
        batch.run(seeds)
        return [batch.edges(k) for k in range(len(segments))]
    finally:
        batch.close()


def pool_counts(tables, n_patches: int) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Sum the integer edge counts of several ray blocks into one sorted edge list (view_factors.py:290-300)."""

    n = int(n_patches)
    keys = np.concatenate([np.asarray(o, np.int64) * n + np.asarray(r, np.int64) for o, r, _ in tables])
    counts = np.concatenate([np.asarray(c, dtype=np.int64) for _, _, c in tables])
    order = np.argsort(keys, kind="stable")
    keys, counts = keys[order], counts[order]
    if keys.size == 0:
        return np.empty(0, np.int32), np.empty(0, np.int32), np.empty(0, np.int64)
    first = np.concatenate(([True], keys[1:] != keys[:-1]))
    starts = np.flatnonzero(first)
    unique = keys[starts]
    return (unique // n).astype(np.int32), (unique % n).astype(np.int32), np.add.reduceat(counts, starts)


def count_sha256(origin: np.ndarray, receiver: np.ndarray, count: np.ndarray, rays_per_origin: np.ndarray) -> str:
    """SHA-256 of a pooled count table, comparable with the case-table hashes (view_factors.py:303-323).

    For origin (int32), receiver (int32), count (int64) and rays per origin
    (int64), in this order: the array name, the NumPy dtype string, the shape
    as int64 and the C-order bytes.
    """

    h = hashlib.sha256()
    parts = (
        ("origin_patch_id", np.int32, origin),
        ("receiver_patch_id", np.int32, receiver),
        ("count", np.int64, count),
        ("rays_per_origin", np.int64, rays_per_origin),
    )
    for name, dtype, values in parts:
        a = np.ascontiguousarray(np.asarray(values, dtype=dtype))
        h.update(name.encode("ascii"))
        h.update(a.dtype.str.encode("ascii"))
        h.update(np.asarray(a.shape, dtype=np.int64).tobytes())
        h.update(a.tobytes())
    return h.hexdigest()
