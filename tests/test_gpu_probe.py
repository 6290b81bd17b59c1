"""The measurement probes of the batched sweep, on a small operator.

prh_probe_tiles splits K into dense 16 x 4 MFMA tiles and a CSR remainder:
tiles + remainder must give the plain product (FP32 tile accumulation of
integer K times FP32 slots: 1e-6 relative), with every nonzero in tiles
(min_fill 1, which exercises the MFMA lane layout on every entry) and with
part of them.  prh_probe_gather runs each slot format and reports a time.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def operator():
    from porerad import _native, exchange

    _native.lib()
    rng = np.random.default_rng(6)
    phases = (rng.random((18, 15, 13)) < 0.15).astype(np.uint8)  # patch count not a multiple of 16
    Q = O.build_surface_patches(phases, 0.01)
    o, r, c = O.first_hit_counts(Q, 96, 17)
    op = exchange.ExchangeOperator(o, r, c, np.full(Q.n_patches, 96, np.int64), Q.patch_area_m2)
    yield op
    op.device.close()


@pytest.mark.parametrize("min_fill", [1, 2, 4])
def test_tiles_plus_remainder_equal_plain_product(operator, min_fill):
    got = operator.device.probe_tiles(min_fill, reps=2)
    assert got["max_rel_difference"] <= 1e-6, got
    if min_fill == 1:
        assert got["remainder_nnz"] == 0 and got["nnz_fraction_in_tiles"] == 1.0
    else:
        assert 0.0 < got["nnz_fraction_in_tiles"] < 1.0
    assert got["plain_ms"] > 0 and got["tiles_ms"] > 0


def test_every_slot_format_sweeps(operator):
    from porerad import _native

    for name, fmt in _native.PROBE_FORMATS.items():
        assert operator.device.probe_gather(fmt, reps=2) > 0.0, name
    with pytest.raises(Exception):
        operator.device.probe_gather(99, reps=1)


def test_reordered_sweep_equals_stored_sweep_reordered(operator):
    rng = np.random.default_rng(21)
    order = rng.permutation(operator.n_patches).astype(np.int32)
    got = operator.device.probe_order(order, reps=2)
    assert got["max_rel_difference"] <= 2e-3, got  # FP16 outputs, sums in another order
    assert min(got["stored_ms"], got["reordered_ms"], got["stored_streamed_ms"]) > 0.0
    with pytest.raises(Exception):
        operator.device.probe_order(np.zeros(operator.n_patches, np.int32), reps=1)


@pytest.mark.parametrize("n_panels", [1, 3])
def test_panel_sweep_equals_plain_sweep(operator, n_panels):
    operator.set_sweep_order(np.random.default_rng(22).permutation(operator.n_patches).astype(np.int32))
    got = operator.device.probe_panels(n_panels, reps=2)
    operator.set_sweep_order(None)
    assert got["max_rel_difference"] <= 2e-3, got  # FP16 outputs, sums in another order
    assert got["plain_ms"] > 0 and got["panels_ms"] > 0 and got["panels_sweep_order_ms"] > 0
