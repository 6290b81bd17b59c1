"""The CPU oracle reproduces the reference's archived integer first-hit counts.

Each of the 12 shipped canonical N=128 orientations (OF/FI/FA/TL x x/y/z) has
two archived ray sets A and B (2 seeds x 64 rays per patch, pooled, R = 128)
whose count table SHA-256 is listed in data/cases_coupled.csv
(raw_count_sha256_A/B, checked by run_coupled.py:78-86).  Six of the 24
hashes run by default; PORERAD_FULL=1 runs all of them.
"""

import numpy as np
import pytest

from golden_data import cases, phases_x
from oracle import oracle as O

DEFAULT = {("OF", "x", "A"), ("FI", "y", "B"), ("FA", "z", "A"), ("TL", "x", "B"), ("TL", "z", "A"), ("FA", "x", "B")}


def _params():
    out = []
    for h in cases()["hash_cases"]:
        for s in "AB":
            marks = [] if (h["structure"], h["axis"], s) in DEFAULT else [pytest.mark.slow]
            out.append(pytest.param(h, s, id=f"{h['structure']}-{h['axis']}-{s}", marks=marks))
    return out


_patch_cache = {}


def _patches(h):
    key = (h["microstructure"], h["axis"])
    if key not in _patch_cache:
        _patch_cache.clear()
        _patch_cache[key] = O.build_surface_patches(phases_x(h["microstructure"], h["axis"]), h["H_m"])
    return _patch_cache[key]


@pytest.mark.parametrize("case,which", _params())
def test_oracle_reproduces_archived_count_hash(case, which):
    import hashlib

    ph = phases_x(case["microstructure"], case["axis"])
    assert hashlib.sha256(ph.tobytes()).hexdigest() == case["phase_sha256_x"]
    P = _patches(case)
    R = case["rays_per_patch_per_block"]
    seeds = case[f"seeds_{which}"]
    tables = [O.first_hit_counts(P, R, s) for s in seeds]
    o, r, c = O.pool_counts(tables, P.n_patches)
    rays = np.full(P.n_patches, R * len(seeds), dtype=np.int64)
    assert O.count_sha256(o, r, c, rays) == case[f"raw_count_sha256_{which}"]
