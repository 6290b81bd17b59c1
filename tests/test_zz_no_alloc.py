"""prh_trace_counts_into (no device allocation, the library's fixed scratch
array) against prh_trace_counts: the same edges, bit for bit, on two calls in
a row, the second smaller than the first.

The caller owns the output arrays; here they come from hipMalloc of the HIP
runtime the library itself is linked against (libamdhip64.so.7), called
through ctypes.  torch is not used: its wheel ships a HIP runtime of its own,
and two runtimes do not share one process."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GUARD = np.int32(np.uint32(0xF9F9F9F9).view(np.int32))  # hipMemset byte 0xF9 in every entry
HIP_MEMCPY_D2H = 2


def _hip():
    from porerad import _native

    _native.lib()  # the library's runtime is loaded first; the same soname resolves to it
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipDeviceSynchronize.argtypes = []
    return hip


class DeviceArrays:
    """Three int32 device arrays of n entries, filled with the guard value."""

    def __init__(self, hip, n):
        self.hip, self.n, self.ptrs = hip, n, []
        for _ in range(3):
            p = ctypes.c_void_p()
            assert hip.hipMalloc(ctypes.byref(p), 4 * n) == 0
            assert hip.hipMemset(p, 0xF9, 4 * n) == 0
            self.ptrs.append(p)
        assert hip.hipDeviceSynchronize() == 0

    def read(self):
        out = []
        for p in self.ptrs:
            a = np.empty(self.n, np.int32)
            assert self.hip.hipMemcpy(a.ctypes.data_as(ctypes.c_void_p), p, 4 * self.n, HIP_MEMCPY_D2H) == 0
            out.append(a)
        return out

    def close(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def _into(P, p0, p1, R, seeds, capacity):
    from porerad import _native

    hip = _hip()
    states = np.ascontiguousarray(np.stack([_native.pcg64_state(s) for s in seeds]), dtype=np.uint64)
    out = DeviceArrays(hip, capacity + 64)  # 64 guard entries past the capacity must stay untouched
    try:
        nnz, reads = ctypes.c_int64(-1), ctypes.c_int64(-1)
        status = _native.lib().prh_trace_counts_into(
            P.geometry().handle, int(p0), int(p1), int(R), len(seeds), _native.ptr(states), out.ptrs[0], out.ptrs[1],
            out.ptrs[2], int(capacity), ctypes.byref(nnz), ctypes.byref(reads))
        return status, nnz.value, reads.value, out.read()
    finally:
        out.close()


def test_trace_counts_into_equals_trace_counts():
    from porerad import _native, synthetic, view_factors

    P = view_factors.build_surface_patches(synthetic.random_fibers(64, 0.10, 3.0, 0), 0.01)
    calls = [(0, P.n_patches, 200, [1, 2]), (1000, 5000, 100, [3])]  # the second call is smaller
    for p0, p1, R, seeds in calls:
        ref = _native.trace_counts(P.geometry(), p0, p1, R, seeds)
        o_ref, r_ref, c_ref = ref.to_host()
        reads_ref = ref.occupancy_reads
        ref.close()
        cap = int(o_ref.size) + 100
        status, nnz, reads, (o, r, c) = _into(P, p0, p1, R, seeds, cap)
        assert status == _native.PRH_OK, _native.last_error()
        assert nnz == o_ref.size and reads == reads_ref
        assert np.array_equal(o[:nnz], o_ref) and np.array_equal(r[:nnz], r_ref)
        assert np.array_equal(c[:nnz].astype(np.int64), c_ref)
        assert np.all(o[nnz:] == GUARD) and np.all(c[nnz:] == GUARD)  # nothing written past the edges


def test_trace_counts_into_limits():
    from porerad import _native, synthetic, view_factors

    P = view_factors.build_surface_patches(synthetic.random_fibers(24, 0.12, 2.0, 0), 0.003)
    ref = _native.trace_counts(P.geometry(), 0, P.n_patches, 50, [4])
    need = ref.nnz
    ref.close()
    # too small an output: nothing past the capacity, the size needed comes back
    status, nnz, _, (o, r, c) = _into(P, 0, P.n_patches, 50, [4], need // 2)
    assert status == _native.PRH_ERR_CAPACITY and nnz == need
    assert np.all(o[need // 2:] == GUARD) and np.all(c[need // 2:] == GUARD)
    # more rays per patch than one pass holds is refused before any launch
    status, _, _, _ = _into(P, 0, 10, 16385, [4], 10)
    assert status == _native.PRH_ERR_CAPACITY
    assert "16384" in _native.last_error()
