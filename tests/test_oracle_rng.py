"""The oracle's SeedSequence/PCG64 restatement against numpy itself (numpy 2.2.6 is
the library the reference draws its rays from, porerad/view_factors.py:149-151)."""

import numpy as np
import pytest

from golden_data import cases
from oracle import oracle as O


def _case_seeds():
    c = cases()
    seeds = set()
    for h in c["hash_cases"]:
        seeds.update(h["seeds_A"] + h["seeds_B"])
    for p in c["prescribed_field"]:
        seeds.update(p["final_seeds"])
    return sorted(seeds)


SPECIAL = [0, 1, 2, 3, 12345, 2**32 - 1, 2**32, 2**63 - 1, 2**64 - 1, 2**64 + 5, 2**128 + 17, 3**200]


@pytest.mark.parametrize("seed", SPECIAL + _case_seeds())
def test_seed_state_matches_numpy(seed):
    st = np.random.PCG64(seed).state["state"]
    ref = [st["state"] >> 64, st["state"] & (2**64 - 1), st["inc"] >> 64, st["inc"] & (2**64 - 1)]
    assert [int(v) for v in O.seed_state(seed)] == ref


def test_known_answer_pcg64_12345():
    # numpy.random.PCG64(12345).state, as printed by numpy 2.2.6
    st = O.seed_state(12345)
    assert (int(st[0]) << 64) | int(st[1]) == 33261208707367790463622745601869196757
    assert (int(st[2]) << 64) | int(st[3]) == 268209174141567072605526753992732310247


@pytest.mark.parametrize("seed", [0, 7, 4734788061198622954])
def test_raw_outputs_match_numpy(seed):
    g = np.random.PCG64(seed)
    assert np.array_equal(O.raw64(seed, 4096), g.random_raw(4096))


@pytest.mark.parametrize("advance", [0, 1, 3, 4 * 64 * 1000, 4 * 128 * 395709 + 12, 2**40 + 3, 2**70 + 11])
def test_advance_then_uniforms_match_numpy(advance):
    seed = 208955243124512515
    g = np.random.PCG64(seed)
    g.advance(advance)
    ref = np.random.Generator(g).random(2048)
    assert np.array_equal(O.uniforms(seed, advance, 2048), ref)
