"""assemble_conduction (porerad/conduction.py:57-97) on the host.

* bit for bit the oracle's assembly (a restatement of the reference's face
  loop with a COO sum) for lateral sizes >= 3, within 2 ulp of the diagonal
  where a lateral size is 1 or 2 (two faces join the same pair of cells);
* a uniform medium conducts the linear profile exactly: A T = b with
  T = T_h + (T_c - T_h)(i + 1/2)/n_x (half-cell wall links);
* symmetric, row sums equal the wall links.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import coupled_oracle as CO


def _field(shape, seed):
    rng = np.random.default_rng(seed)
    return np.where(rng.random(shape) < 0.3, 0.2, 0.026) * rng.uniform(0.5, 2.0, shape)


@pytest.mark.parametrize("shape", [(5, 4, 3), (7, 6, 5), (16, 3, 9), (3, 3, 3), (12, 10, 11), (1, 4, 4)])
def test_matches_oracle_bitwise(shape):
    from porerad import conduction

    k = _field(shape, 1)
    got = conduction.assemble_conduction(k, 0.013)
    ref = CO.Conduction(k, 0.013)
    A, B = got.matrix, ref.A
    assert np.array_equal(A.indptr, B.indptr) and np.array_equal(A.indices, B.indices)
    assert np.array_equal(A.data.view(np.int64), B.data.view(np.int64))
    assert np.array_equal(got.hot_rhs, ref.hot) and np.array_equal(got.cold_rhs, ref.cold)


@pytest.mark.parametrize("shape", [(4, 2, 3), (3, 1, 5), (6, 1, 1), (9, 8, 2), (2, 5, 2)])
def test_thin_grids_match_oracle(shape):
    from porerad import conduction

    k = _field(shape, 2)
    A = conduction.assemble_conduction(k, 0.01).matrix
    B = CO.Conduction(k, 0.01).A
    assert np.array_equal(A.indptr, B.indptr) and np.array_equal(A.indices, B.indices)
    assert np.max(np.abs(A.data - B.data)) <= 4.5e-16 * np.max(np.abs(B.data))


def test_uniform_medium_conducts_the_linear_profile():
    from porerad import conduction

    nx, ny, nz, H, Th, Tc = 10, 4, 5, 0.02, 700.0, 300.0
    sysm = conduction.assemble_conduction(np.full((nx, ny, nz), 0.7), H)
    T = Th + (Tc - Th) * ((np.arange(nx) + 0.5) / nx)
    T = np.broadcast_to(T[:, None, None], (nx, ny, nz)).reshape(-1)
    b = sysm.boundary_rhs(Th, Tc)
    assert np.max(np.abs(sysm.matrix @ T - b)) <= 1e-12 * np.max(np.abs(b))
    q_hot, q_cold = sysm.wall_fluxes(T, Th, Tc)
    assert abs(q_hot - 0.7 * (Th - Tc) / H) <= 1e-9 * q_hot and abs(q_cold - q_hot) <= 1e-9 * q_hot


def test_symmetric_and_row_sums_are_wall_links():
    from porerad import conduction

    k = _field((6, 5, 4), 3)
    s = conduction.assemble_conduction(k, 0.01)
    A = s.matrix
    assert abs(A - A.T).max() == 0.0
    rows = np.asarray(A.sum(axis=1)).ravel()
    assert np.allclose(rows, s.hot_rhs + s.cold_rhs, rtol=0, atol=1e-15 * np.max(np.abs(A.data)))
