"""The library's sin/cos (prh_libm.h) against the host C library, on the host.

generate_rays (porerad/view_factors.py:168-176) takes NumPy's cos and sin of
phi = 2 pi u, u = k 2^-53, which are the C library's.  The same source runs
in the ray generation on the device; here its host build is compared bit for
bit with numpy (and numpy with math, i.e. the C library) on 10^8 azimuths of
that form, on every azimuth 2 pi m 2^-24 of the Monte Carlo draws, and on
dense runs of consecutive doubles around every branch point of the method.
No GPU is needed.
"""

from __future__ import annotations

import math

import numpy as np

TWO_PI = 2.0 * math.pi


def _same(x):
    from porerad import _native

    s, c = _native.libm_sincos(x, on_device=False)
    assert np.array_equal(s.view(np.int64), np.sin(x).view(np.int64))
    assert np.array_equal(c.view(np.int64), np.cos(x).view(np.int64))


def test_numpy_sin_cos_are_the_c_library():
    rng = np.random.default_rng(3)
    x = TWO_PI * rng.random(20000)
    assert all(math.sin(v) == s for v, s in zip(x, np.sin(x)))
    assert all(math.cos(v) == c for v, c in zip(x, np.cos(x)))


def test_generate_rays_azimuths_1e8():
    rng = np.random.default_rng(20261204)
    for _ in range(10):
        k = rng.integers(0, 2**53, size=10_000_000, dtype=np.uint64)
        u = k.astype(np.float64) * (1.0 / 9007199254740992.0)
        _same(TWO_PI * u)


def test_monte_carlo_azimuths_all():
    m = np.arange(2**24, dtype=np.float64)
    _same(TWO_PI * (m / 16777216.0))


def test_branch_points_dense():
    centres = [2.0**-27, 2.0**-26, 0.126, 0.855469, float.fromhex("0x1.b6p-1"), 2.426265, float.fromhex("0x1.368fdp1"), math.pi / 4, math.pi / 2,
               3 * math.pi / 4, math.pi, 5 * math.pi / 4, 3 * math.pi / 2, 7 * math.pi / 4, TWO_PI]
    for c in centres:
        bits = np.float64(c).view(np.int64) + np.arange(-200_000, 200_000, dtype=np.int64)
        x = bits.view(np.float64)
        _same(np.concatenate([x[x > 0], -x[x > 0]]))


def test_other_arguments():
    rng = np.random.default_rng(5)
    _same(np.concatenate([rng.uniform(-1e8, 1e8, 1_000_000), rng.uniform(-50, 50, 1_000_000),
                          np.array([0.0, -0.0, 1e-300, -1e-300, 5e-324])]))


def test_rejects_out_of_domain():
    import pytest

    from porerad import _native

    for bad in (np.inf, np.nan, 2e8):
        with pytest.raises(ValueError):
            _native.libm_sincos(np.array([1.0, bad]), on_device=False)
