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


def _numpy_rays(P, p0, p1, R, seed):
    """generate_rays (porerad/view_factors.py:147-182) evaluated with NumPy's own
    routines: Generator.random, np.clip, np.remainder, np.sqrt, np.cos, np.sin and
    np.linalg.norm, vectorised over the patch range."""

    bitgen = np.random.PCG64(seed)
    bitgen.advance(p0 * R * 4)
    u = np.random.Generator(bitgen).random((p1 - p0, R, 4))
    normal = P.patch_normal[p0:p1]
    axis = np.argmax(normal != 0.0, axis=1)
    sign = normal[np.arange(p1 - p0), axis]
    tang = np.clip(u[:, :, :2], 1e-12, 1.0 - 1e-12) - 0.5
    origin = np.repeat(P.patch_origin[p0:p1, None, :], R, axis=1)
    lateral = np.array([[1, 2], [0, 2], [0, 1]])[axis]
    rows = np.arange(p1 - p0)
    for t in range(2):
        origin[rows, :, lateral[:, t]] += tang[:, :, t]
    origin[:, :, 1] = np.remainder(origin[:, :, 1], P.solid.shape[1])
    origin[:, :, 2] = np.remainder(origin[:, :, 2], P.solid.shape[2])
    r = np.sqrt(u[:, :, 2])
    phi = 2.0 * np.pi * u[:, :, 3]
    d = np.empty((p1 - p0, R, 3))
    d[rows, :, axis] = sign[:, None] * np.sqrt(1.0 - u[:, :, 2])
    d[rows, :, lateral[:, 0]] = r * np.cos(phi)
    d[rows, :, lateral[:, 1]] = sign[:, None] * (r * np.sin(phi))
    d = d.reshape(-1, 3)
    d /= np.linalg.norm(d, axis=1)[:, None]
    return origin.reshape(-1, 3), d


def test_oracle_rays_equal_numpy_bit_for_bit():
    # The oracle's C cos/sin must be the separate calls np.cos/np.sin make
    # (a fused sincos differs from them in about 0.1 % of arguments).
    rng = np.random.default_rng(12)
    solid = (rng.random((24, 20, 22)) < 0.15).astype(np.uint8)
    P = O.build_surface_patches(solid, 0.01)
    for p0, p1, seed in ((0, P.n_patches, 4734788061198622954), (37, 411, 3)):
        o_ref, d_ref = _numpy_rays(P, p0, p1, 200, seed)
        o, d = O.generate_rays(P, p0, p1, 200, seed)
        assert np.array_equal(o.view(np.int64), o_ref.view(np.int64))
        assert np.array_equal(d.view(np.int64), d_ref.view(np.int64))
