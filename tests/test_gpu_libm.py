"""The device build of the library's sin/cos against the host C library.

The ray generation (porerad/view_factors.py:168-176) and the Monte Carlo
Lambertian azimuth (porerad/properties.py:217-219) run prh_libm.h on the
device; its results must be NumPy's (the C library's) bit for bit.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TWO_PI = 2.0 * math.pi


def _same_on_device(x):
    from porerad import _native

    s, c = _native.libm_sincos(x, on_device=True)
    assert np.array_equal(s.view(np.int64), np.sin(x).view(np.int64))
    assert np.array_equal(c.view(np.int64), np.cos(x).view(np.int64))


def test_device_generate_rays_azimuths():
    rng = np.random.default_rng(77)
    for _ in range(2):
        k = rng.integers(0, 2**53, size=10_000_000, dtype=np.uint64)
        _same_on_device(TWO_PI * (k.astype(np.float64) * (1.0 / 9007199254740992.0)))


def test_device_monte_carlo_azimuths_all():
    m = np.arange(2**24, dtype=np.float64)
    _same_on_device(TWO_PI * (m / 16777216.0))


def test_device_branch_points_and_other_arguments():
    parts = []
    for c in [2.0**-27, 2.0**-26, 0.126, 0.855469, 2.426265, math.pi / 2, math.pi, 3 * math.pi / 2, TWO_PI]:
        bits = np.float64(c).view(np.int64) + np.arange(-100_000, 100_000, dtype=np.int64)
        x = bits.view(np.float64)
        parts += [x[x > 0], -x[x > 0]]
    rng = np.random.default_rng(9)
    parts += [rng.uniform(-1e8, 1e8, 500_000), np.array([0.0, -0.0, 5e-324])]
    _same_on_device(np.concatenate(parts))
