"""Access to the committed golden fixtures (tests/golden/)."""

from __future__ import annotations

import functools
import hashlib
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
AXES = {"x": 0, "y": 1, "z": 2}


@functools.lru_cache(maxsize=None)
def cases() -> dict:
    with open(os.path.join(GOLDEN, "cases.json")) as handle:
        return json.load(handle)


@functools.lru_cache(maxsize=None)
def _archive():
    with np.load(os.path.join(GOLDEN, "geometries.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def phases(name: str) -> np.ndarray:
    """The shipped phase array (uint8, original orientation), integrity-checked."""

    a = _archive()
    shape = tuple(int(s) for s in a[f"{name}__shape"])
    arr = np.unpackbits(a[f"{name}__bits"])[: int(np.prod(shape))].reshape(shape).astype(np.uint8)
    if hashlib.sha256(arr.tobytes()).digest() != a[f"{name}__sha256"].tobytes():
        raise ValueError(f"fixture {name} is corrupt")
    return arr


def phases_x(name: str, axis: str) -> np.ndarray:
    return np.ascontiguousarray(np.moveaxis(phases(name), AXES[axis], 0))


@functools.lru_cache(maxsize=None)
def configs() -> dict:
    """CPU-oracle results of the full-size benchmark workloads (tests/golden/make_config_fixtures.py)."""

    with open(os.path.join(GOLDEN, "configs.json")) as handle:
        return json.load(handle)
