"""Independent checks of the oracle tracer (porerad/view_factors.py:70-258 restated).

* Closed-form patch tables of tiny grids.
* Brute force: for rays without near-ties, the first solid face or wall
  crossed, found with exact rational arithmetic on the FP64 inputs, equals
  the oracle's receiver.
* An empty box sends every wall ray to the opposite wall.
"""

from fractions import Fraction
import math

import numpy as np
import pytest

from oracle import oracle as O


def test_single_voxel_patch_table():
    phases = np.zeros((5, 5, 5), np.uint8)
    phases[2, 2, 2] = 1
    P = O.build_surface_patches(phases, 0.005)
    assert P.n_wall == 50 and P.n_patches == 56
    assert list(P.patch_class[:4]) == [0, 1, 0, 1]
    assert list(P.patch_face[50:]) == [0, 1, 2, 3, 4, 5]
    cell = (2 * 5 + 2) * 5 + 2
    assert np.all(P.patch_cell[50:] == cell)
    off = 0.5 + 1.0e-9
    assert P.patch_origin[50].tolist() == [2.5 + off, 2.5, 2.5]
    assert P.patch_origin[51].tolist() == [2.5 - off, 2.5, 2.5]
    assert P.patch_origin[54].tolist() == [2.5, 2.5, 2.5 + off]
    assert np.array_equal(P.cell_face_patch[cell], np.arange(50, 56))
    assert P.patch_origin[0].tolist() == [1.0e-9, 0.5, 0.5]
    assert P.patch_origin[1].tolist() == [5.0 - 1.0e-9, 0.5, 0.5]
    assert np.allclose(P.patch_area_m2, (0.005 / 5) ** 2)


def test_boundary_solid_faces_are_not_patches_and_lateral_faces_wrap():
    phases = np.zeros((3, 4, 4), np.uint8)
    phases[0, 0, 0] = 1  # on the hot boundary and on the periodic y/z seam
    P = O.build_surface_patches(phases, 0.003)
    # no hot-wall patch at (0, 0), 16 cold ones, 15 hot ones
    assert P.hot_wall_patch[0, 0] == -1 and P.n_wall == 31
    faces = P.patch_face[P.n_wall:].tolist()
    assert faces == [0, 2, 3, 4, 5]  # no -x face on the boundary
    # the -y face origin wraps to y = 4 - (0.5 + 1e-9) + 0.5 mod 4
    o = P.patch_origin[P.n_wall + 2]
    assert o[1] == (0.5 - (0.5 + 1.0e-9)) % 4.0


def _exact_first_hit(P, o, d):
    """Receiver by exact rational walk, or None when a near-tie makes it ambiguous."""

    nx, ny, nz = P.solid.shape
    O3 = [Fraction(float(v)) for v in o]
    D3 = [Fraction(float(v)) for v in d]
    if D3[0] == 0:
        return None
    cell = [math.floor(v) for v in O3]
    for a in range(3):
        if abs(float(O3[a]) - round(float(O3[a]))) < 1e-9:
            return None
    step = [1 if v > 0 else (-1 if v < 0 else 0) for v in D3]
    nxt = []
    for a in range(3):
        if step[a] == 0:
            nxt.append(None)
        else:
            plane = cell[a] + 1 if step[a] > 0 else cell[a]
            nxt.append((plane - O3[a]) / D3[a])
    if P.solid[cell[0], cell[1] % ny, cell[2] % nz]:
        return None
    for _ in range(10000):
        times = [(t, a) for a, t in enumerate(nxt) if t is not None]
        times.sort()
        t, a = times[0]
        if len(times) > 1 and float(times[1][0] - t) < 1e-9:
            return None
        cell[a] += step[a]
        nxt[a] = nxt[a] + abs(1 / D3[a])
        if cell[0] < 0 or cell[0] >= nx:
            y = O3[1] + D3[1] * t
            z = O3[2] + D3[2] * t
            if abs(float(y) - round(float(y))) < 1e-9 or abs(float(z) - round(float(z))) < 1e-9:
                return None
            jj, kk = math.floor(y) % ny, math.floor(z) % nz
            return int((P.hot_wall_patch if cell[0] < 0 else P.cold_wall_patch)[jj, kk])
        jj, kk = cell[1] % ny, cell[2] % nz
        if P.solid[cell[0], jj, kk]:
            face = (1 if step[0] > 0 else 0) if a == 0 else ((3 if step[1] > 0 else 2) if a == 1 else (5 if step[2] > 0 else 4))
            return int(P.cell_face_patch[(cell[0] * ny + jj) * nz + kk, face])
    return None


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_oracle_tracer_matches_exact_rational_walk(seed):
    rng = np.random.default_rng(100 + seed)
    phases = (rng.random((6, 5, 7)) < 0.18).astype(np.uint8)
    P = O.build_surface_patches(phases, 0.006)
    o, d = O.generate_rays(P, 0, P.n_patches, 6, 900 + seed)
    rec, _ = O.trace_rays(P, o, d)
    checked = 0
    for m in range(0, d.shape[0], 3):
        exact = _exact_first_hit(P, o[m], d[m])
        if exact is None:
            continue
        checked += 1
        assert rec[m] == exact, (m, o[m], d[m])
    assert checked > 0.8 * len(range(0, d.shape[0], 3))


def test_empty_box_sends_every_wall_ray_across():
    P = O.build_surface_patches(np.zeros((8, 6, 6), np.uint8), 0.008)
    assert P.n_patches == P.n_wall == 72
    o, r, c = O.first_hit_counts(P, 50, 3)
    assert np.all(np.bincount(o, weights=c, minlength=72) == 50)
    assert np.all(P.patch_class[o] != P.patch_class[r])


def test_unresolved_rays_raise():
    phases = np.zeros((4, 4, 4), np.uint8)
    phases[1, 1, 1] = 1
    P = O.build_surface_patches(phases, 0.004)
    o = np.array([[0.5, 0.5, 0.5], [1.5, 1.5, 1.5]])
    d = np.array([[0.0, 1.0, 0.0], [1.0, 0.0, 0.0]])  # dx == 0; start inside solid
    with pytest.raises(RuntimeError, match="2 rays did not resolve a receiving patch"):
        O.trace_first_hit(P, o, d)
