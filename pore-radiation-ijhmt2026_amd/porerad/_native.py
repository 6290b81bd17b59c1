"""ctypes binding of the porerad HIP library (``libporerad_hip.so``, C ABI in
``include/porerad_hip.h``).

The library is the only compute path of this package: there is no CPU
fallback.  If the shared library is missing or cannot be loaded, every entry
point raises ``RuntimeError`` naming the cause.  Status codes map to the
exception types of the reference package (porerad/view_factors.py:268-269,
porerad/exchange.py:70-71, 243-244).
"""

from __future__ import annotations

import ctypes
import os
import threading

import numpy as np

# The in-tree build; PORERAD_HIP_LIB names another copy of the library.
LIB_PATH = os.environ.get("PORERAD_HIP_LIB") or os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                                             "libporerad_hip.so")

PRH_OK = 0
PRH_ERR_ARG = 1
PRH_ERR_HIP = 2
PRH_ERR_UNRESOLVED = 3
PRH_ERR_COUNTS = 4
PRH_ERR_SERIES = 5
PRH_ERR_CAPACITY = 6
PRH_ERR_BALANCE = 7

APPLY_SYMMETRIC, APPLY_MATVEC, APPLY_RMATVEC, APPLY_RAW, APPLY_RAW_T = 0, 1, 2, 3, 4
FP64, MIXED, FP64_BATCHED = 0, 1, 2
# slot formats of prh_probe_gather
PROBE_FORMATS = {"fp32x32": 0, "fp16x64": 1, "fp16x32": 2, "fp64x16": 3, "fp64x32": 4}
BATCH_SLOTS = 64
BATCH_DIRECT = 1

# Every function the C ABI exports: name -> (argtypes, restype).
_P = ctypes.c_void_p
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64
_F64 = ctypes.c_double
SIGNATURES = {
    "prh_last_error": ([], ctypes.c_char_p),
    "prh_version": ([], ctypes.c_char_p),
    "prh_last_timings": ([_P, _I32], ctypes.c_int),
    "prh_pcg64_seed": ([_P, _I32, _P], ctypes.c_int),
    "prh_libm_sincos": ([_P, _I64, _I32, _P, _P], ctypes.c_int),
    "prh_tracer_set_variant": ([_I32], ctypes.c_int),
    "prh_tracer_variants": ([_P, _P, _P, _P], ctypes.c_int),
    "prh_tracer_stats": ([_P, _P, _I32], ctypes.c_int),
    "prh_tracer_set_grain": ([_I32], ctypes.c_int),
    "prh_tracer_grain": ([_P], ctypes.c_int),
    "prh_tracer_trim": ([], ctypes.c_int),
    "prh_tracer_set_workspace_limit": ([_I64], ctypes.c_int),
    "prh_tracer_workspace": ([_P, _P], ctypes.c_int),
    "prh_geometry_create": ([_P, _I32, _I32, _I32, _F64, _P], ctypes.c_int),
    "prh_geometry_create_occupancy": ([_P, _I32, _I32, _I32, _P], ctypes.c_int),
    "prh_geometry_sizes": ([_P, _P, _P], ctypes.c_int),
    "prh_geometry_patch_table": ([_P, _P, _P, _P, _P], ctypes.c_int),
    "prh_geometry_destroy": ([_P], None),
    "prh_mc_properties": ([_P, _I64, ctypes.c_uint32, _F64, _I32, _I32, _P, _P, _P, _P], ctypes.c_int),
    "prh_mc_penetration": ([_P, _I64, ctypes.c_uint32, _F64, _I32, _I32, _P], ctypes.c_int),
    "prh_phase_field_create": ([_P, _I32, _I32, _I32, _P], ctypes.c_int),
    "prh_phase_field_read": ([_P, _P], ctypes.c_int),
    "prh_phase_field_destroy": ([_P], None),
    "prh_scalar_counts": ([_P, _I32, _P], ctypes.c_int),
    "prh_periodic_chords": ([_P, _I32, _P, _I64, _P, _P, _P, _I32], ctypes.c_int),
    "prh_phase_field_fill": ([_P, ctypes.c_uint8], ctypes.c_int),
    "prh_phase_field_copy": ([_P, _P], ctypes.c_int),
    "prh_phase_field_count": ([_P, ctypes.c_uint8, _P], ctypes.c_int),
    "prh_stamp_shapes": ([_P, _I32, _P, _P], ctypes.c_int),
    "prh_stamp_offsets": ([_P, _I32, _P, _I32, _P, _P], ctypes.c_int),
    "prh_stamp_long_cylinders": ([_P, _I32, _P, _P, _P, _P, _P], ctypes.c_int),
    "prh_threshold_field": ([_P, _P, _P, _F64, _P], ctypes.c_int),
    "prh_generate_rays": ([_P, _I64, _I64, _I32, _P, _P, _P], ctypes.c_int),
    "prh_trace_rays": ([_P, _P, _P, _I64, _P, _P], ctypes.c_int),
    "prh_trace_counts": ([_P, _I64, _I64, _I32, _I32, _P, _P, _P, _P], ctypes.c_int),
    "prh_trace_counts_into": ([_P, _I64, _I64, _I32, _I32, _P, _P, _P, _P, _I64, _P, _P], ctypes.c_int),
    "prh_edges_copy": ([_P, _P, _P, _P], ctypes.c_int),
    "prh_edges_equal": ([_P, _P, _P], ctypes.c_int),
    "prh_trace_batch_create": ([_I32, _P, _P, _P, _P, _P, _I64, _I64, _I32, _I32, _P], ctypes.c_int),
    "prh_trace_batch_run": ([_P, _P, _P, _P, _P], ctypes.c_int),
    "prh_trace_batch_edges": ([_P, _I32, _P, _P, _P], ctypes.c_int),
    "prh_trace_batch_edges_device": ([_P, _I32, _P], ctypes.c_int),
    "prh_trace_batch_info": ([_P, _P, _I32], ctypes.c_int),
    "prh_trace_batch_destroy": ([_P], None),
    "prh_edges_destroy": ([_P], None),
    "prh_operator_create": ([_I64, _P, _P, _P, _I64, _P, _P, _F64, _P], ctypes.c_int),
    "prh_operator_set_scale": ([_P, _P], ctypes.c_int),
    "prh_operator_balance": ([_P, _F64, _P, _P], ctypes.c_int),
    "prh_operator_get_scale": ([_P, _P], ctypes.c_int),
    "prh_operator_create_device": ([_P, _I64, _P, _P, _F64, _P], ctypes.c_int),
    "prh_operator_edges": ([_P, _P, _P, _P], ctypes.c_int),
    "prh_operator_nnz": ([_P, _P, _P], ctypes.c_int),
    "prh_operator_csr": ([_P, _P, _P, _P], ctypes.c_int),
    "prh_operator_reciprocity": ([_P, _P], ctypes.c_int),
    "prh_operator_stats": ([_P, _P, _I32], ctypes.c_int),
    "prh_operator_set_row_order": ([_P, _P], ctypes.c_int),
    "prh_probe_gather": ([_P, _I32, _I32, _P], ctypes.c_int),
    "prh_probe_tiles": ([_P, _I32, _I32, _P, _I32], ctypes.c_int),
    "prh_probe_order": ([_P, _P, _I32, _P, _I32], ctypes.c_int),
    "prh_probe_panels": ([_P, _I32, _I32, _P, _I32], ctypes.c_int),
    "prh_batch_stats": ([_P, _P, _I32], ctypes.c_int),
    "prh_operator_apply": ([_P, _I32, _P, _P], ctypes.c_int),
    "prh_radiosity": ([_P, _I32, _P, _P, _I32, _F64, _I32, _I32, _P, _P, _P], ctypes.c_int),
    "prh_operator_destroy": ([_P], None),
    "prh_batch_create": ([_P, _P], ctypes.c_int),
    "prh_batch_bytes": ([_P, _P], ctypes.c_int),
    "prh_batch_free_slots": ([_P, _P], ctypes.c_int),
    "prh_batch_admit": ([_P, _P, _P, _I32, _F64, _P], ctypes.c_int),
    "prh_batch_set_field": ([_P, _P, _P, _I32, _F64], ctypes.c_int),
    "prh_batch_admit_linear": ([_P, _F64, _F64, _F64, _F64, _P], ctypes.c_int),
    "prh_batch_step": ([_P, _I32, _P, _P], ctypes.c_int),
    "prh_batch_channels": ([_P, _I32, _P, _P, _P], ctypes.c_int),
    "prh_batch_release": ([_P, _I32], ctypes.c_int),
    "prh_batch_result": ([_P, _I32, _P, _P, _P], ctypes.c_int),
    "prh_batch_destroy": ([_P], None),
    "prh_coupled_create": ([_P, _P, _I64, _P, _P, _P, _P, _P, _I32, _P], ctypes.c_int),
    "prh_coupled_destroy": ([_P], None),
    "prh_coupled_set": ([_P, _I32, _P], ctypes.c_int),
    "prh_coupled_get": ([_P, _I32, _P], ctypes.c_int),
    "prh_coupled_fill": ([_P, _I32, _F64], ctypes.c_int),
    "prh_coupled_axpby": ([_P, _I32, _F64, _I32, _F64], ctypes.c_int),
    "prh_coupled_divide": ([_P, _I32, _I32, _F64], ctypes.c_int),
    "prh_coupled_dot": ([_P, _I32, _I32, _P], ctypes.c_int),
    "prh_coupled_stats": ([_P, _I32, _P], ctypes.c_int),
    "prh_coupled_matvec": ([_P, _I32, _I32], ctypes.c_int),
    "prh_coupled_radiation": ([_P, _I32, _F64, _F64, _F64, _F64, _I32, _P, _P], ctypes.c_int),
    "prh_coupled_residual": ([_P, _I32, _I32, _F64, _F64, _I32], ctypes.c_int),
    "prh_coupled_linearize": ([_P, _I32, _F64, _F64], ctypes.c_int),
    "prh_coupled_jacobian": ([_P, _I32, _F64, _F64, _I32, _P], ctypes.c_int),
    "prh_coupled_set_preconditioner": ([_P, _I32, _F64, _I32], ctypes.c_int),
    "prh_coupled_set_precision": ([_P, _I32], ctypes.c_int),
    "prh_coupled_precondition": ([_P, _I32, _I32], ctypes.c_int),
    "prh_coupled_orthogonalize": ([_P, _I32, _I32, _I32, _P], ctypes.c_int),
    "prh_coupled_combine": ([_P, _I32, _I32, _I32, _I32, _P], ctypes.c_int),
    "prh_coupled_cg": ([_P, _I32, _I32, _F64, _I32, _P, _P], ctypes.c_int),
    "prh_coupled_state": ([_P, _P, _P, _P], ctypes.c_int),
}

_lock = threading.Lock()
_lib = None
_load_error: str | None = None


def load(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load the library and declare every exported signature (no GPU call is made)."""

    L = ctypes.CDLL(path)
    for name, (argtypes, restype) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.argtypes = argtypes
        fn.restype = restype
    return L


def lib() -> ctypes.CDLL:
    global _lib, _load_error
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(
                    f"porerad HIP library not built: {LIB_PATH} is missing "
                    "(run __graft_entry__.build() or pore-radiation-ijhmt2026_amd/build.py)")
            try:
                _lib = load(LIB_PATH)
            except OSError as exc:  # pragma: no cover - depends on the ROCm install
                _load_error = str(exc)
                raise RuntimeError(f"porerad HIP library failed to load: {exc}") from exc
    return _lib


def last_error() -> str:
    msg = lib().prh_last_error()
    return msg.decode() if msg else ""


def check(status: int, what: str = "") -> None:
    """Raise the reference's exception type for a non-zero status."""

    if status == PRH_OK:
        return
    message = last_error() or f"{what} failed with status {status}"
    if status in (PRH_ERR_UNRESOLVED, PRH_ERR_SERIES, PRH_ERR_BALANCE):
        raise RuntimeError(message)
    if status in (PRH_ERR_COUNTS, PRH_ERR_ARG, PRH_ERR_CAPACITY):
        raise ValueError(message)
    raise RuntimeError(f"HIP error in {what}: {message}")


def ptr(a: np.ndarray) -> ctypes.c_void_p:
    return a.ctypes.data_as(ctypes.c_void_p)


def seed_words(seed: int) -> np.ndarray:
    """Little-endian uint32 words of a non-negative integer (SeedSequence entropy)."""

    value = int(seed)
    if value < 0:
        raise ValueError("expected non-negative integer")
    words = [0] if value == 0 else []
    while value > 0:
        words.append(value & 0xFFFFFFFF)
        value >>= 32
    return np.asarray(words, dtype=np.uint32)


def pcg64_state(seed: int) -> np.ndarray:
    """{state_hi, state_lo, inc_hi, inc_lo} of numpy.random.PCG64(seed)."""

    words = seed_words(seed)
    out = np.zeros(4, dtype=np.uint64)
    check(lib().prh_pcg64_seed(ptr(words), int(words.size), ptr(out)), "prh_pcg64_seed")
    return out


def libm_sincos(x: np.ndarray, on_device: bool) -> tuple[np.ndarray, np.ndarray]:
    """sin(x), cos(x) by the library's restatement of the C library's sin/cos
    (the routine the ray generation runs), evaluated on the host or the device."""

    x = np.ascontiguousarray(x, dtype=np.float64)
    s = np.empty_like(x)
    c = np.empty_like(x)
    check(lib().prh_libm_sincos(ptr(x), int(x.size), int(bool(on_device)), ptr(s), ptr(c)), "prh_libm_sincos")
    return s, c


def last_timings() -> np.ndarray:
    """[generation, traversal, counting, compaction] ms of the last prh_trace_counts call."""

    out = np.zeros(4, dtype=np.float64)
    check(lib().prh_last_timings(ptr(out), 4), "prh_last_timings")
    return out


def tracer_variants() -> list[dict]:
    n = ctypes.c_int32()
    cur = ctypes.c_int32()
    check(lib().prh_tracer_variants(ctypes.byref(n), ctypes.byref(cur), None, None), "prh_tracer_variants")
    refill = np.zeros(n.value, np.int32)
    occ = np.zeros(n.value, np.int32)
    check(lib().prh_tracer_variants(None, None, ptr(refill), ptr(occ)), "prh_tracer_variants")
    return [{"variant": v, "refill_min": int(refill[v]), "occupancy": int(occ[v]), "current": v == cur.value}
            for v in range(n.value)]


def tracer_stats(reset: bool = False) -> dict:
    """Wave-steps and refills of the lean traversal since the last reset."""

    steps = ctypes.c_int64()
    refills = ctypes.c_int64()
    check(lib().prh_tracer_stats(ctypes.byref(steps), ctypes.byref(refills), 1 if reset else 0), "prh_tracer_stats")
    return {"wave_steps": int(steps.value), "refills": int(refills.value)}


def set_tracer_variant(variant: int) -> None:
    check(lib().prh_tracer_set_variant(int(variant)), "prh_tracer_set_variant")


def tracer_grain() -> int:
    g = ctypes.c_int32()
    check(lib().prh_tracer_grain(ctypes.byref(g)), "prh_tracer_grain")
    return int(g.value)


def set_tracer_grain(records: int) -> None:
    check(lib().prh_tracer_set_grain(int(records)), "prh_tracer_set_grain")


def trim_tracer() -> None:
    """Free the tracer's grow-only device workspace."""

    check(lib().prh_tracer_trim(), "prh_tracer_trim")


def set_tracer_workspace_limit(nbytes: int) -> None:
    """Bound the tracer's device workspace; larger inputs are traced in several passes."""

    check(lib().prh_tracer_set_workspace_limit(int(nbytes)), "prh_tracer_set_workspace_limit")


def tracer_workspace() -> tuple[int, int]:
    """(bytes held on the current device, bound) of the tracer's workspace."""

    held, limit = ctypes.c_int64(), ctypes.c_int64()
    check(lib().prh_tracer_workspace(ctypes.byref(held), ctypes.byref(limit)), "prh_tracer_workspace")
    return int(held.value), int(limit.value)


class Geometry:
    """Owner of a device patch table (``prh_geometry``)."""

    def __init__(self, solid_x: np.ndarray, H_m: float) -> None:
        solid = np.ascontiguousarray(solid_x, dtype=np.uint8)
        if solid.ndim != 3:
            raise ValueError("expected a 3-D phase array")
        self.shape = tuple(int(s) for s in solid.shape)
        handle = ctypes.c_void_p()
        L = lib()
        check(L.prh_geometry_create(ptr(solid), *self.shape, float(H_m), ctypes.byref(handle)),
              "prh_geometry_create")
        self._lib = L
        self.handle = handle
        n, nw = ctypes.c_int64(), ctypes.c_int64()
        check(L.prh_geometry_sizes(handle, ctypes.byref(n), ctypes.byref(nw)), "prh_geometry_sizes")
        self.n_patches, self.n_wall = int(n.value), int(nw.value)

    @classmethod
    def occupancy_only(cls, solid_x: np.ndarray) -> "Geometry":
        """The occupancy bitsets of an x-oriented solid array, without patch tables (the Monte Carlo's input)."""

        solid = np.ascontiguousarray(solid_x, dtype=np.uint8)
        if solid.ndim != 3:
            raise ValueError("expected a 3-D phase array")
        self = cls.__new__(cls)
        self.shape = tuple(int(s) for s in solid.shape)
        handle = ctypes.c_void_p()
        L = lib()
        check(L.prh_geometry_create_occupancy(ptr(solid), *self.shape, ctypes.byref(handle)),
              "prh_geometry_create_occupancy")
        self._lib = L
        self.handle = handle
        self.n_patches, self.n_wall = 0, 0
        return self

    def patch_table(self):
        n = self.n_patches
        cls = np.empty(n, np.int8)
        cell = np.empty(n, np.int32)
        face = np.empty(n, np.int8)
        origin = np.empty((n, 3), np.float64)
        check(self._lib.prh_geometry_patch_table(self.handle, ptr(cls), ptr(cell), ptr(face), ptr(origin)),
              "prh_geometry_patch_table")
        return cls, cell, face, origin

    def close(self) -> None:
        if getattr(self, "handle", None) is not None and self.handle.value:
            self._lib.prh_geometry_destroy(self.handle)
            self.handle = None

    def __del__(self):  # pragma: no cover - interpreter shutdown order
        try:
            self.close()
        except Exception:
            pass


DESC_CELLS, DESC_SOLID, DESC_PORE = 0, 1, 2
DESC_BOX_FACES, DESC_PORE_RUNS, DESC_PERIODIC_FACES = 3, 6, 9
DESC_LARGEST_CLUSTER, DESC_CLUSTERS, DESC_PERCOLATES = 12, 13, 14
DESC_COUNT = 15


class PhaseField:
    """Owner of a device copy of a phase array (``prh_phase_field``)."""

    def __init__(self, phases: np.ndarray | None = None, shape: tuple[int, int, int] | None = None) -> None:
        """A copy of ``phases``, or a field of zeros of ``shape``."""

        if phases is not None:
            p = np.ascontiguousarray(phases, dtype=np.uint8)
            if p.ndim != 3:
                raise ValueError("expected a 3-D phase array")
            self.shape = tuple(int(s) for s in p.shape)
            data = ptr(p)
        else:
            self.shape = tuple(int(s) for s in shape)
            data = None
        handle = ctypes.c_void_p()
        L = lib()
        check(L.prh_phase_field_create(data, *self.shape, ctypes.byref(handle)), "prh_phase_field_create")
        self._lib = L
        self.handle = handle

    def fill(self, value: int) -> None:
        check(self._lib.prh_phase_field_fill(self.handle, int(value)), "prh_phase_field_fill")

    def copy_from(self, other: "PhaseField") -> None:
        check(self._lib.prh_phase_field_copy(self.handle, other.handle), "prh_phase_field_copy")

    def count(self, value: int) -> int:
        n = ctypes.c_int64()
        check(self._lib.prh_phase_field_count(self.handle, int(value), ctypes.byref(n)), "prh_phase_field_count")
        return int(n.value)

    def stamp_shapes(self, shapes: np.ndarray, kinds: np.ndarray) -> None:
        """OR short shapes into the field (rows of 16 doubles, see prh_stamp_shapes)."""

        s = np.ascontiguousarray(shapes, np.float64).reshape(-1, 16)
        k = np.ascontiguousarray(kinds, np.int32)
        if len(k) != len(s):
            raise ValueError("one kind per shape")
        check(self._lib.prh_stamp_shapes(self.handle, len(k), ptr(s), ptr(k)), "prh_stamp_shapes")

    def stamp_offsets(self, nodes: np.ndarray, offsets: np.ndarray, shift=(0, 0, 0)) -> None:
        n = np.ascontiguousarray(nodes, np.int32).reshape(-1, 3)
        o = np.ascontiguousarray(offsets, np.int32).reshape(-1, 3)
        sh = np.ascontiguousarray(shift, np.int32)
        check(self._lib.prh_stamp_offsets(self.handle, len(n), ptr(n), len(o), ptr(o), ptr(sh)), "prh_stamp_offsets")

    def stamp_long_cylinders(self, nodes: list, rq_t: np.ndarray, radius: np.ndarray, shift: np.ndarray) -> None:
        """Stamp several long cylinders (see prh_stamp_long_cylinders); nodes is a list of (n_c, 3) arrays."""

        counts = np.asarray([len(n) for n in nodes], np.int64)
        start = np.zeros(len(nodes) + 1, np.int32)
        start[1:] = np.cumsum(counts)
        flat = np.ascontiguousarray(np.concatenate(nodes) if nodes else np.zeros((0, 3)), np.int32)
        rq = np.ascontiguousarray(rq_t, np.float64)
        r = np.ascontiguousarray(radius, np.float64)
        sh = np.ascontiguousarray(shift, np.int32)
        check(self._lib.prh_stamp_long_cylinders(self.handle, len(nodes), ptr(start), ptr(flat), ptr(rq), ptr(r),
                                                 ptr(sh)), "prh_stamp_long_cylinders")

    def threshold(self, field: np.ndarray, q: float, noise: np.ndarray | None = None) -> float:
        """Cells >= np.quantile(values, q) become 1, the rest 0; values = field (+ smoothed noise)."""

        f = np.ascontiguousarray(field, np.float64)
        nz = None if noise is None else np.ascontiguousarray(noise, np.float64)
        if f.shape != self.shape or (nz is not None and nz.shape != self.shape):
            raise ValueError("field and noise must have the shape of the phase field")
        out = ctypes.c_double()
        check(self._lib.prh_threshold_field(self.handle, ptr(f), None if nz is None else ptr(nz), float(q),
                                            ctypes.byref(out)), "prh_threshold_field")
        return float(out.value)

    def read(self) -> np.ndarray:
        out = np.empty(self.shape, np.uint8)
        check(self._lib.prh_phase_field_read(self.handle, ptr(out)), "prh_phase_field_read")
        return out

    def scalar_counts(self, connectivity: bool = True) -> np.ndarray:
        """The int64 counts of prh_scalar_counts (indices DESC_*)."""

        out = np.zeros(DESC_COUNT, np.int64)
        check(self._lib.prh_scalar_counts(self.handle, 1 if connectivity else 0, ptr(out)), "prh_scalar_counts")
        return out

    def periodic_chords(self, axis: int, hist_len: int):
        """(chords int32 in the reference's order, number of all-pore lines, histogram int64[hist_len])."""

        length = self.shape[axis]
        lines = int(np.prod(self.shape)) // length
        capacity = lines * ((length + 1) // 2)
        chords = np.empty(max(capacity, 1), np.int32)
        n, allp = ctypes.c_int64(), ctypes.c_int64()
        hist = np.zeros(int(hist_len), np.int64)
        check(self._lib.prh_periodic_chords(self.handle, int(axis), ptr(chords), int(capacity), ctypes.byref(n),
                                            ctypes.byref(allp), ptr(hist), int(hist_len)), "prh_periodic_chords")
        return chords[: n.value].copy(), int(allp.value), hist

    def close(self) -> None:
        if getattr(self, "handle", None) is not None and self.handle.value:
            self._lib.prh_phase_field_destroy(self.handle)
            self.handle = None

    def __enter__(self) -> "PhaseField":
        return self

    def __exit__(self, *exc) -> None:
        self.close()

    def __del__(self):  # pragma: no cover - interpreter shutdown order
        try:
            self.close()
        except Exception:
            pass


class Edges:
    """Owner of a device-resident edge table (``prh_edges``)."""

    def __init__(self, handle: ctypes.c_void_p, nnz: int, occupancy_reads: int) -> None:
        self._lib = lib()
        self.handle = handle
        self.nnz = int(nnz)
        self.occupancy_reads = int(occupancy_reads)

    def to_host(self):
        origin = np.empty(self.nnz, np.int32)
        receiver = np.empty(self.nnz, np.int32)
        count = np.empty(self.nnz, np.int64)
        check(self._lib.prh_edges_copy(self.handle, ptr(origin), ptr(receiver), ptr(count)), "prh_edges_copy")
        return origin, receiver, count

    def equals(self, other: "Edges") -> bool:
        """Whether two device edge tables hold the same entries (compared on the device)."""

        eq = ctypes.c_int32()
        check(self._lib.prh_edges_equal(self.handle, other.handle, ctypes.byref(eq)), "prh_edges_equal")
        return bool(eq.value)

    def close(self) -> None:
        if getattr(self, "handle", None) is not None and self.handle.value:
            self._lib.prh_edges_destroy(self.handle)
            self.handle = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass


def trace_counts(geometry: Geometry, p0: int, p1: int, rays_per_patch: int, seeds) -> Edges:
    """Fused generate + trace + count of patches [p0, p1) over one or more seeds."""

    states = np.ascontiguousarray(np.stack([pcg64_state(s) for s in seeds]), dtype=np.uint64)
    handle = ctypes.c_void_p()
    nnz = ctypes.c_int64()
    reads = ctypes.c_int64()
    check(lib().prh_trace_counts(geometry.handle, int(p0), int(p1), int(rays_per_patch), int(len(seeds)),
                                 ptr(states), ctypes.byref(handle), ctypes.byref(nnz), ctypes.byref(reads)),
          "prh_trace_counts")
    return Edges(handle, nnz.value, reads.value)


class TraceBatch:
    """Owner of a graph-captured multi-geometry trace pass (``prh_trace_batch``).

    ``segments`` is a list of ``(geometry, p0, p1, rays_per_patch, n_seeds)``;
    every :meth:`run` replays the pass with one seed list per segment and
    leaves each segment's pooled edge table on the device.
    """

    INFO = ("device_bytes", "chunks", "pieces", "ray_capacity", "edge_capacity", "rays", "streams", "graph",
            "variant")

    def __init__(self, segments, ray_capacity: int = 0, edge_capacity: int = 0, streams: int = 0,
                 graph: bool = True) -> None:
        self._lib = lib()
        self.geometries = [seg[0] for seg in segments]  # kept alive with the batch
        n = len(segments)
        handles = (ctypes.c_void_p * n)(*[g.handle.value for g in self.geometries])
        p0 = np.array([seg[1] for seg in segments], np.int64)
        p1 = np.array([seg[2] for seg in segments], np.int64)
        rays = np.array([seg[3] for seg in segments], np.int32)
        self.n_seeds = np.array([seg[4] for seg in segments], np.int32)
        handle = ctypes.c_void_p()
        check(self._lib.prh_trace_batch_create(n, ctypes.cast(handles, ctypes.c_void_p), ptr(p0), ptr(p1), ptr(rays),
                                               ptr(self.n_seeds), int(ray_capacity), int(edge_capacity),
                                               int(streams), 0 if graph else BATCH_DIRECT, ctypes.byref(handle)),
              "prh_trace_batch_create")
        self.handle = handle
        self.n = n
        self.nnz = np.zeros(n, np.int64)
        self.occupancy_reads = np.zeros(n, np.int64)
        self.elapsed_ms = 0.0
        info = np.zeros(len(self.INFO), np.int64)
        check(self._lib.prh_trace_batch_info(handle, ptr(info), len(self.INFO)), "prh_trace_batch_info")
        self.info = dict(zip(self.INFO, (int(v) for v in info)))

    def run(self, seeds) -> None:
        """Replay the pass; ``seeds[k]`` lists segment k's ``n_seeds[k]`` seeds."""

        if len(seeds) != self.n:
            raise ValueError(f"expected seeds for {self.n} segments")
        rows = []
        for k, block in enumerate(seeds):
            if len(block) != int(self.n_seeds[k]):
                raise ValueError(f"segment {k} takes {int(self.n_seeds[k])} seeds")
            rows.extend(pcg64_state(s) for s in block)
        states = np.ascontiguousarray(np.stack(rows), dtype=np.uint64)
        ms = ctypes.c_double()
        check(self._lib.prh_trace_batch_run(self.handle, ptr(states), ptr(self.nnz), ptr(self.occupancy_reads),
                                            ctypes.byref(ms)), "prh_trace_batch_run")
        self.elapsed_ms = float(ms.value)

    def edges(self, segment: int):
        """(origin, receiver, count) of a segment from the last run."""

        m = int(self.nnz[segment])
        origin, receiver, count = np.empty(m, np.int32), np.empty(m, np.int32), np.empty(m, np.int64)
        check(self._lib.prh_trace_batch_edges(self.handle, int(segment), ptr(origin), ptr(receiver), ptr(count)),
              "prh_trace_batch_edges")
        return origin, receiver, count

    def edge_table(self, segment: int) -> "Edges":
        """A segment's edges of the last run as a device edge table (no host copy)."""

        handle = ctypes.c_void_p()
        check(self._lib.prh_trace_batch_edges_device(self.handle, int(segment), ctypes.byref(handle)),
              "prh_trace_batch_edges_device")
        return Edges(handle, int(self.nnz[segment]), int(self.occupancy_reads[segment]))

    def close(self) -> None:
        if getattr(self, "handle", None) is not None and self.handle.value:
            self._lib.prh_trace_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass


class Operator:
    """Owner of a device exchange operator (``prh_operator``)."""

    @classmethod
    def from_edges(cls, n: int, edges: "Edges", rays, area, delta: float) -> "Operator":
        """Assemble on the device from a device edge table (``prh_operator_create_device``)."""

        self = cls.__new__(cls)
        self._lib = lib()
        self.n = int(n)
        rays = np.ascontiguousarray(rays, dtype=np.int64)
        area = np.ascontiguousarray(area, dtype=np.float64)
        if rays.shape != (self.n,) or area.shape != (self.n,):
            raise ValueError(f"expected {self.n} ray counts and areas")
        handle = ctypes.c_void_p()
        check(self._lib.prh_operator_create_device(edges.handle, self.n, ptr(rays), ptr(area), float(delta),
                                                   ctypes.byref(handle)), "prh_operator_create_device")
        self.handle = handle
        a, b = ctypes.c_int64(), ctypes.c_int64()
        check(self._lib.prh_operator_nnz(handle, ctypes.byref(a), ctypes.byref(b)), "prh_operator_nnz")
        self.nnz_symmetric, self.nnz_raw = int(a.value), int(b.value)
        return self

    def reciprocity(self) -> float:
        """Largest relative reciprocity defect over the count table's pairs (computed on the device)."""

        err = ctypes.c_double()
        check(self._lib.prh_operator_reciprocity(self.handle, ctypes.byref(err)), "prh_operator_reciprocity")
        return float(err.value)

    def set_row_order(self, order: np.ndarray | None) -> None:
        """Sweep order of the batched engine's rows (a permutation of the patches), or None."""

        if order is None:
            check(self._lib.prh_operator_set_row_order(self.handle, None), "prh_operator_set_row_order")
            return
        order = np.ascontiguousarray(order, dtype=np.int32)
        if order.shape != (self.n,):
            raise ValueError("order must hold one patch index per patch")
        check(self._lib.prh_operator_set_row_order(self.handle, ptr(order)), "prh_operator_set_row_order")

    def probe_gather(self, fmt: int, reps: int = 20) -> float:
        """Device ms of one F_p^T sweep over every slot in slot format ``fmt`` (a measurement probe)."""

        ms = ctypes.c_double()
        check(self._lib.prh_probe_gather(self.handle, int(fmt), int(reps), ctypes.byref(ms)), "prh_probe_gather")
        return float(ms.value)

    def probe_tiles(self, min_fill: int, reps: int = 20) -> dict:
        """Dense 16 x 4 MFMA tiles (blocks with >= min_fill nonzeros) + CSR remainder vs the plain gather."""

        out = np.zeros(7)
        check(self._lib.prh_probe_tiles(self.handle, int(min_fill), int(reps), ptr(out), 7), "prh_probe_tiles")
        return {"plain_ms": float(out[0]), "tiles_ms": float(out[1]), "remainder_ms": float(out[2]),
                "max_rel_difference": float(out[3]), "tiles": int(out[4]), "nnz_fraction_in_tiles": float(out[5]),
                "remainder_nnz": int(out[6])}

    def probe_order(self, order: np.ndarray, reps: int = 20) -> dict:
        """FP16 x 64 sweep in the stored patch order and renumbered by ``order``, plain and streamed."""

        order = np.ascontiguousarray(order, dtype=np.int32)
        if order.shape != (self.n,):
            raise ValueError("order must hold one patch index per patch")
        out = np.zeros(5)
        check(self._lib.prh_probe_order(self.handle, ptr(order), int(reps), ptr(out), 5), "prh_probe_order")
        return {"stored_ms": float(out[0]), "stored_streamed_ms": float(out[1]), "reordered_ms": float(out[2]),
                "reordered_streamed_ms": float(out[3]), "max_rel_difference": float(out[4])}

    def probe_panels(self, n_panels: int, reps: int = 20) -> dict:
        """FP16 x 64 sweep in column panels against the plain sweep (a measurement probe)."""

        out = np.zeros(4)
        check(self._lib.prh_probe_panels(self.handle, int(n_panels), int(reps), ptr(out), 4), "prh_probe_panels")
        return {"plain_ms": float(out[0]), "panels_ms": float(out[1]), "panels_sweep_order_ms": float(out[2]),
                "max_rel_difference": float(out[3])}

    def csr(self):
        """(rowptr, col, val) of the symmetric FP64 matrix S held on the device."""

        rowptr = np.empty(self.n + 1, np.int64)
        col = np.empty(self.nnz_symmetric, np.int32)
        val = np.empty(self.nnz_symmetric, np.float64)
        check(self._lib.prh_operator_csr(self.handle, ptr(rowptr), ptr(col), ptr(val)), "prh_operator_csr")
        return rowptr, col, val

    def edges(self):
        """(origin, receiver, count) the operator was built from."""

        m = self.nnz_raw
        origin, receiver, count = np.empty(m, np.int32), np.empty(m, np.int32), np.empty(m, np.int64)
        check(self._lib.prh_operator_edges(self.handle, ptr(origin), ptr(receiver), ptr(count)), "prh_operator_edges")
        return origin, receiver, count

    def __init__(self, n: int, origin, receiver, count, rays, area, delta: float) -> None:
        self._lib = lib()
        self.n = int(n)
        origin = np.ascontiguousarray(origin, dtype=np.int32)
        receiver = np.ascontiguousarray(receiver, dtype=np.int32)
        count = np.ascontiguousarray(count, dtype=np.int64)
        rays = np.ascontiguousarray(rays, dtype=np.int64)
        area = np.ascontiguousarray(area, dtype=np.float64)
        handle = ctypes.c_void_p()
        check(self._lib.prh_operator_create(self.n, ptr(origin), ptr(receiver), ptr(count), int(count.size),
                                            ptr(rays), ptr(area), float(delta), ctypes.byref(handle)),
              "prh_operator_create")
        self.handle = handle
        a, b = ctypes.c_int64(), ctypes.c_int64()
        check(self._lib.prh_operator_nnz(handle, ctypes.byref(a), ctypes.byref(b)), "prh_operator_nnz")
        self.nnz_symmetric, self.nnz_raw = int(a.value), int(b.value)

    def stats(self) -> dict:
        """Device time of the FP64 reflection series run on this operator."""

        out = np.zeros(9)
        check(self._lib.prh_operator_stats(self.handle, ptr(out), 9), "prh_operator_stats")
        return {"series": int(out[0]), "reflections": int(out[1]), "ms": float(out[2]), "nnz": int(out[3]),
                "n": int(out[4]), "integer_form": bool(out[5]), "batched_passes": int(out[6]),
                "batched_slot_passes": int(out[7]), "batched_ms": float(out[8])}

    def balance(self, tolerance: float) -> tuple[int, int]:
        """Balance on the device (method 0 none, 1 Anderson, 2 multiplicative; iterations)."""

        method, iterations = ctypes.c_int32(), ctypes.c_int32()
        check(self._lib.prh_operator_balance(self.handle, float(tolerance), ctypes.byref(method),
                                             ctypes.byref(iterations)), "prh_operator_balance")
        return int(method.value), int(iterations.value)

    def get_scale(self) -> np.ndarray:
        out = np.empty(self.n, np.float64)
        check(self._lib.prh_operator_get_scale(self.handle, ptr(out)), "prh_operator_get_scale")
        return out

    def set_scale(self, scale: np.ndarray) -> None:
        scale = np.ascontiguousarray(scale, dtype=np.float64)
        check(self._lib.prh_operator_set_scale(self.handle, ptr(scale)), "prh_operator_set_scale")

    def apply(self, mode: int, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float64)
        if x.shape != (self.n,):
            raise ValueError(f"expected a vector of {self.n} values")
        y = np.empty(self.n, np.float64)
        check(self._lib.prh_operator_apply(self.handle, int(mode), ptr(x), ptr(y)), "prh_operator_apply")
        return y

    def radiosity(self, emitted: np.ndarray, emissivity: np.ndarray, relative_tolerance: float,
                  max_reflections: int, precision: int = FP64):
        E = np.ascontiguousarray(np.atleast_2d(emitted), dtype=np.float64)
        n_rhs = E.shape[0]
        eps = np.ascontiguousarray(emissivity, dtype=np.float64)
        per_rhs = 1 if eps.ndim == 2 else 0
        absorbed = np.empty_like(E)
        residual = np.empty(n_rhs, np.float64)
        iterations = np.empty(n_rhs, np.int32)
        check(self._lib.prh_radiosity(self.handle, n_rhs, ptr(E), ptr(eps), per_rhs, float(relative_tolerance),
                                      int(max_reflections), int(precision), ptr(absorbed), ptr(residual),
                                      ptr(iterations)), "prh_radiosity")
        return absorbed, residual, iterations

    def close(self) -> None:
        if getattr(self, "handle", None) is not None and self.handle.value:
            self._lib.prh_operator_destroy(self.handle)
            self.handle = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Owner of a batched radiosity engine (``prh_batch``) bound to an Operator."""

    def __init__(self, operator: Operator) -> None:
        self._lib = lib()
        self.operator = operator
        self.n = operator.n
        handle = ctypes.c_void_p()
        check(self._lib.prh_batch_create(operator.handle, ctypes.byref(handle)), "prh_batch_create")
        self.handle = handle

    @property
    def bytes(self) -> int:
        """Device bytes the engine holds (slot vectors, partial sums, staging, field)."""

        b = ctypes.c_int64()
        check(self._lib.prh_batch_bytes(self.handle, ctypes.byref(b)), "prh_batch_bytes")
        return int(b.value)

    def stats(self) -> dict:
        """Device time of the engine's ticks (HIP events around each tick)."""

        out = np.zeros(6)
        check(self._lib.prh_batch_stats(self.handle, ptr(out), 6), "prh_batch_stats")
        return {"ticks": int(out[0]), "ms": float(out[1]), "slot_ticks": int(out[2]), "nnz": int(out[3]),
                "n": int(out[4]), "value_bytes": int(out[5])}

    def free_slots(self) -> int:
        m = ctypes.c_uint64()
        check(self._lib.prh_batch_free_slots(self.handle, ctypes.byref(m)), "prh_batch_free_slots")
        return int(m.value)

    def admit(self, emitted: np.ndarray, emissivity, relative_tolerance: float) -> int:
        E = np.ascontiguousarray(emitted, dtype=np.float64)
        eps = np.ascontiguousarray(np.atleast_1d(np.asarray(emissivity, dtype=np.float64)))
        per_patch = 1 if eps.size == self.n else 0
        if E.shape != (self.n,) or (per_patch == 0 and eps.size != 1):
            raise ValueError("emission and emissivity must have one value per patch")
        slot = ctypes.c_int32()
        check(self._lib.prh_batch_admit(self.handle, ptr(E), ptr(eps), per_patch, float(relative_tolerance),
                                        ctypes.byref(slot)), "prh_batch_admit")
        return int(slot.value)

    def set_field(self, patch_class: np.ndarray, cell_x: np.ndarray, nx: int, cross_section_m2: float) -> None:
        cls = np.ascontiguousarray(patch_class, dtype=np.int8)
        cx = np.ascontiguousarray(cell_x, dtype=np.int32)
        if cls.shape != (self.n,) or cx.shape != (self.n,):
            raise ValueError("patch_class and cell_x must have one value per patch")
        check(self._lib.prh_batch_set_field(self.handle, ptr(cls), ptr(cx), int(nx), float(cross_section_m2)),
              "prh_batch_set_field")

    def admit_linear(self, T_hot_K: float, T_cold_K: float, emissivity: float, relative_tolerance: float) -> int:
        slot = ctypes.c_int32()
        check(self._lib.prh_batch_admit_linear(self.handle, float(T_hot_K), float(T_cold_K), float(emissivity),
                                               float(relative_tolerance), ctypes.byref(slot)),
              "prh_batch_admit_linear")
        return int(slot.value)

    def step(self, max_ticks: int = 1) -> tuple[int, int]:
        """Run up to max_ticks sweeps (returns early when a slot finishes): (done mask, active mask)."""

        done, active = ctypes.c_uint64(), ctypes.c_uint64()
        check(self._lib.prh_batch_step(self.handle, int(max_ticks), ctypes.byref(done), ctypes.byref(active)),
              "prh_batch_step")
        return int(done.value), int(active.value)

    def run_until_done(self) -> int:
        """Step until at least one slot has finished (or nothing runs); the done mask."""

        while True:
            done, active = self.step(64)
            if done or not active:
                return done

    def channels(self, slot: int) -> tuple[float, float, float, float, int]:
        """(q_h, q_c, q_i, residual power, sweeps) of a finished linear-field slot; keeps the slot."""

        q = np.empty(3, np.float64)
        residual = ctypes.c_double()
        iterations = ctypes.c_int32()
        check(self._lib.prh_batch_channels(self.handle, int(slot), ptr(q), ctypes.byref(residual),
                                           ctypes.byref(iterations)), "prh_batch_channels")
        return float(q[0]), float(q[1]), float(q[2]), float(residual.value), int(iterations.value)

    def result(self, slot: int):
        absorbed = np.empty(self.n, np.float64)
        residual = ctypes.c_double()
        iterations = ctypes.c_int32()
        check(self._lib.prh_batch_result(self.handle, int(slot), ptr(absorbed), ctypes.byref(residual),
                                         ctypes.byref(iterations)), "prh_batch_result")
        return absorbed, float(residual.value), int(iterations.value)

    def release(self, slot: int) -> None:
        check(self._lib.prh_batch_release(self.handle, int(slot)), "prh_batch_release")

    def close(self) -> None:
        if getattr(self, "handle", None) is not None and self.handle.value:
            self._lib.prh_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass


class Coupled:
    """Owner of the device state of a coupled solve (``prh_coupled``): cell-vector slots + conduction matrix.

    ``geometry`` and ``operator`` (both or neither) enable the radiation
    operations; without them the handle serves the pure-conduction solve.
    """

    def __init__(self, matrix, hot_rhs, cold_rhs, n_vectors: int, geometry: Geometry | None = None,
                 operator: Operator | None = None) -> None:
        self._lib = lib()
        rowptr = np.ascontiguousarray(matrix.indptr, dtype=np.int64)
        col = np.ascontiguousarray(matrix.indices, dtype=np.int32)
        val = np.ascontiguousarray(matrix.data, dtype=np.float64)
        hot = np.ascontiguousarray(hot_rhs, dtype=np.float64)
        cold = np.ascontiguousarray(cold_rhs, dtype=np.float64)
        self.n = int(matrix.shape[0])
        if hot.shape != (self.n,) or cold.shape != (self.n,):
            raise ValueError("wall response vectors must have one value per cell")
        self.n_patches = operator.n if operator is not None else 0
        handle = ctypes.c_void_p()
        check(self._lib.prh_coupled_create(geometry.handle if geometry is not None else None,
                                           operator.handle if operator is not None else None, self.n, ptr(rowptr),
                                           ptr(col), ptr(val), ptr(hot), ptr(cold), int(n_vectors),
                                           ctypes.byref(handle)), "prh_coupled_create")
        self.handle = handle
        self.geometry, self.operator = geometry, operator  # kept alive while the handle uses them

    def set(self, v: int, x: np.ndarray) -> None:
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
        if x.size != self.n:
            raise ValueError(f"expected {self.n} values")
        check(self._lib.prh_coupled_set(self.handle, int(v), ptr(x)), "prh_coupled_set")

    def get(self, v: int) -> np.ndarray:
        out = np.empty(self.n, np.float64)
        check(self._lib.prh_coupled_get(self.handle, int(v), ptr(out)), "prh_coupled_get")
        return out

    def fill(self, v: int, value: float) -> None:
        check(self._lib.prh_coupled_fill(self.handle, int(v), float(value)), "prh_coupled_fill")

    def axpby(self, y: int, a: float, x: int, b: float) -> None:
        check(self._lib.prh_coupled_axpby(self.handle, int(y), float(a), int(x), float(b)), "prh_coupled_axpby")

    def divide(self, y: int, x: int, d: float) -> None:
        check(self._lib.prh_coupled_divide(self.handle, int(y), int(x), float(d)), "prh_coupled_divide")

    def dot(self, x: int, y: int) -> float:
        out = ctypes.c_double()
        check(self._lib.prh_coupled_dot(self.handle, int(x), int(y), ctypes.byref(out)), "prh_coupled_dot")
        return float(out.value)

    def stats(self, x: int) -> tuple[float, float, float, float, int]:
        """(||x||_2, ||x||_inf, min, max, non-finite count)."""

        out = np.zeros(5, np.float64)
        check(self._lib.prh_coupled_stats(self.handle, int(x), ptr(out)), "prh_coupled_stats")
        return float(out[0]), float(out[1]), float(out[2]), float(out[3]), int(out[4])

    def matvec(self, x: int, y: int) -> None:
        check(self._lib.prh_coupled_matvec(self.handle, int(x), int(y)), "prh_coupled_matvec")

    def radiation(self, T: int, T_hot: float, T_cold: float, eps: float, rtol: float, Q: int) -> tuple[float, int]:
        res, its = ctypes.c_double(), ctypes.c_int32()
        check(self._lib.prh_coupled_radiation(self.handle, int(T), float(T_hot), float(T_cold), float(eps),
                                              float(rtol), int(Q), ctypes.byref(res), ctypes.byref(its)),
              "prh_coupled_radiation")
        return float(res.value), int(its.value)

    def residual(self, T: int, Q: int, T_hot: float, T_cold: float, R: int) -> None:
        check(self._lib.prh_coupled_residual(self.handle, int(T), int(Q), float(T_hot), float(T_cold), int(R)),
              "prh_coupled_residual")

    def linearize(self, T: int, T_hot: float, T_cold: float) -> None:
        check(self._lib.prh_coupled_linearize(self.handle, int(T), float(T_hot), float(T_cold)),
              "prh_coupled_linearize")

    def jacobian(self, v: int, eps: float, rtol: float, out: int) -> int:
        its = ctypes.c_int32()
        check(self._lib.prh_coupled_jacobian(self.handle, int(v), float(eps), float(rtol), int(out),
                                             ctypes.byref(its)), "prh_coupled_jacobian")
        return int(its.value)

    def set_precision(self, precision: str) -> None:
        """Radiosity solves of the Jacobian-vector products: "fp64" (series) or "mixed" (engine)."""

        code = {"fp64": FP64, "mixed": MIXED}.get(precision)
        if code is None:
            raise ValueError("precision must be 'fp64' or 'mixed'")
        check(self._lib.prh_coupled_set_precision(self.handle, code), "prh_coupled_set_precision")

    def set_preconditioner(self, T: int, eps: float, steps: int) -> None:
        check(self._lib.prh_coupled_set_preconditioner(self.handle, int(T), float(eps), int(steps)),
              "prh_coupled_set_preconditioner")

    def precondition(self, rhs: int, out: int) -> None:
        check(self._lib.prh_coupled_precondition(self.handle, int(rhs), int(out)), "prh_coupled_precondition")

    def orthogonalize(self, basis: int, count: int, w: int) -> np.ndarray:
        h = np.zeros(count + 1, np.float64)
        check(self._lib.prh_coupled_orthogonalize(self.handle, int(basis), int(count), int(w), ptr(h)),
              "prh_coupled_orthogonalize")
        return h

    def combine(self, out: int, x: int, basis: int, y: np.ndarray) -> None:
        y = np.ascontiguousarray(y, dtype=np.float64)
        check(self._lib.prh_coupled_combine(self.handle, int(out), int(x), int(basis), int(y.size), ptr(y)),
              "prh_coupled_combine")

    def cg(self, b: int, x: int, rtol: float, max_iterations: int) -> tuple[int, float]:
        its, rel = ctypes.c_int32(), ctypes.c_double()
        check(self._lib.prh_coupled_cg(self.handle, int(b), int(x), float(rtol), int(max_iterations),
                                       ctypes.byref(its), ctypes.byref(rel)), "prh_coupled_cg")
        return int(its.value), float(rel.value)

    def state(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """(patch temperatures, emitted, absorbed) of the last radiation evaluation."""

        Tp, E, A = (np.empty(self.n_patches, np.float64) for _ in range(3))
        check(self._lib.prh_coupled_state(self.handle, ptr(Tp), ptr(E), ptr(A)), "prh_coupled_state")
        return Tp, E, A

    def close(self) -> None:
        if getattr(self, "handle", None) is not None and self.handle.value:
            self._lib.prh_coupled_destroy(self.handle)
            self.handle = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass
