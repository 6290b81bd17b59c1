"""Host utilities of porerad.coupled (coupled.py:45-73, 232-315, 450-490)
against the oracle's restatement of the reference and against closed forms.

* flexible_gmres (Givens-rotation FGMRES) solves nonsymmetric systems to its
  tolerance with fixed and with varying preconditioners, through restarts,
  and converges in the same number of steps as the oracle's (lstsq-based)
  FGMRES to within one;
* FixedJacobiPCG equals the oracle's fixed-step Jacobi-PCG to round-off and,
  with as many steps as unknowns, solves a small SPD system;
* the patch/cell maps equal the oracle's on a random medium;
* wall_flux_metrics gives the oracle's fluxes and the reference's keys, and
  acceptance_failures names each failed rule.
"""

from __future__ import annotations

import types

import numpy as np
import pytest
from scipy import sparse

from oracle import coupled_oracle as CO
from oracle import oracle as O


def _system(n, seed, nonsym=0.3):
    rng = np.random.default_rng(seed)
    A = sparse.random(n, n, density=0.05, random_state=seed, format="csr") * nonsym
    A = A + sparse.diags(2.0 + rng.random(n))
    return A.tocsr(), rng.normal(size=n)


@pytest.mark.parametrize("restart,seed", [(30, 0), (5, 1), (3, 2)])
def test_flexible_gmres_solves_and_matches_oracle_steps(restart, seed):
    from porerad import coupled

    A, b = _system(300, seed)
    inv = 1.0 / A.diagonal()
    res = coupled.flexible_gmres(lambda v: A @ v, b, lambda v: inv * v, restart=restart, max_iterations=90,
                                 relative_tolerance=1e-10, absolute_tolerance=1e-13)
    x_ref, ok_ref, steps_ref = CO.fgmres(lambda v: A @ v, b, lambda v: inv * v, restart=restart,
                                         max_iterations=90, rtol=1e-10, atol=1e-13)
    exact = sparse.linalg.spsolve(A.tocsc(), b)
    assert res.converged and ok_ref
    assert np.linalg.norm(A @ res.solution - b) <= 1e-10 * np.linalg.norm(b)
    assert np.max(np.abs(res.solution - exact)) <= 1e-8 * np.max(np.abs(exact))
    assert abs(res.iterations - steps_ref) <= 1
    assert res.true_residual_relative <= 1e-10


def test_flexible_gmres_with_a_varying_preconditioner():
    from porerad import coupled

    A, b = _system(200, 4, nonsym=0.5)
    inv = 1.0 / A.diagonal()
    calls = [0]

    def varying(v):  # a different (still good) approximate inverse on every call
        calls[0] += 1
        return inv * v * (1.0 + 0.2 * np.sin(calls[0]))

    res = coupled.flexible_gmres(lambda v: A @ v, b, varying, restart=20, max_iterations=90,
                                 relative_tolerance=1e-9, absolute_tolerance=0.0)
    assert res.converged
    assert np.linalg.norm(A @ res.solution - b) <= 1e-9 * np.linalg.norm(b)


def test_flexible_gmres_reports_non_convergence():
    from porerad import coupled

    A, b = _system(200, 5, nonsym=3.0)
    res = coupled.flexible_gmres(lambda v: A @ v, b, lambda v: v, restart=2, max_iterations=4,
                                 relative_tolerance=1e-14, absolute_tolerance=0.0)
    assert not res.converged and res.iterations == 4
    assert res.true_residual_relative == pytest.approx(np.linalg.norm(A @ res.solution - b) / np.linalg.norm(b))


def test_fixed_jacobi_pcg_matches_oracle_and_solves_small_systems():
    from porerad import conduction, coupled

    rng = np.random.default_rng(7)
    k = np.where(rng.random((6, 5, 4)) < 0.3, 0.2, 0.026)
    system = conduction.assemble_conduction(k, 0.01)
    shift = rng.uniform(0.0, 1e-4, k.size)
    rhs = rng.normal(size=k.size)
    for steps in (1, 10, 20):
        got = coupled.FixedJacobiPCG(system, shift, steps)(rhs)
        ref = CO.jacobi_pcg(system.matrix, shift, steps)(rhs)
        assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))
    A = sparse.csr_matrix(np.array([[4.0, 1.0, 0.0], [1.0, 3.0, 1.0], [0.0, 1.0, 2.0]]))
    small = types.SimpleNamespace(matrix=A)
    x = coupled.FixedJacobiPCG(small, np.zeros(3), 3)(np.array([1.0, 2.0, 3.0]))
    assert np.allclose(A @ x, [1.0, 2.0, 3.0], rtol=0, atol=1e-12)


def _patches(seed=3):
    from porerad import view_factors as vf

    rng = np.random.default_rng(seed)
    Q = O.build_surface_patches((rng.random((10, 9, 8)) < 0.2).astype(np.uint8), 0.01)
    P = vf.SurfacePatches(Q.solid, Q.patch_class, Q.patch_cell, Q.patch_face, Q.patch_origin, Q.patch_normal,
                          Q.patch_area_m2, Q.cell_face_patch, Q.hot_wall_patch, Q.cold_wall_patch, Q.H_m)
    return P, Q, rng


def test_patch_cell_maps_match_oracle():
    from porerad import coupled

    P, Q, rng = _patches()
    T = rng.uniform(300.0, 700.0, P.solid.shape)
    power = rng.normal(size=P.n_patches)
    assert np.array_equal(coupled.patch_temperature(P, T, 700.0, 300.0), CO.patch_temperature(Q, T, 700.0, 300.0))
    assert np.array_equal(coupled.cell_power(P, power).ravel(), np.asarray(CO.cell_power(Q, power)).ravel())
    assert np.array_equal(coupled.emission_diagonal(P, T, 0.6), CO.emission_diagonal(Q, T, 0.6))


def test_wall_flux_metrics_and_acceptance():
    from porerad import conduction, coupled

    P, Q, rng = _patches(5)
    k = np.where(P.solid == 1, 0.2, 0.026)
    system = conduction.assemble_conduction(k, P.H_m)
    T = np.linspace(700.0, 300.0, P.solid.shape[0])[:, None, None] * np.ones(P.solid.shape)
    emitted, absorbed = rng.uniform(0, 1e-3, P.n_patches), rng.uniform(0, 1e-3, P.n_patches)
    radiosity = types.SimpleNamespace(global_energy_residual_W=1e-12)
    state = types.SimpleNamespace(patch_net_power_W=absorbed - emitted, radiosity=radiosity)
    result = types.SimpleNamespace(temperature_K=T.ravel(), state=state, converged=True, linear_residuals_pass=True)
    m = coupled.wall_flux_metrics(system, P, result, 700.0, 300.0)
    cond = CO.Conduction(k, P.H_m)
    qh, rh, qc, rc = CO.wall_fluxes(cond, Q, T.ravel(), (None, emitted, absorbed, None, None), 700.0, 300.0)
    assert list(m)[:6] == ["q_cond_hot_W_m2", "q_rad_hot_W_m2", "q_total_hot_W_m2", "q_cond_cold_W_m2",
                           "q_rad_cold_W_m2", "q_total_cold_W_m2"]
    assert (m["q_cond_hot_W_m2"], m["q_rad_hot_W_m2"], m["q_cond_cold_W_m2"], m["q_rad_cold_W_m2"]) == (qh, rh, qc, rc)
    assert m["q_total_hot_W_m2"] == qh + rh and m["temperature_bound_violations"] == 0
    assert m["k_eff_apparent_W_mK"] == (qh + rh) * P.H_m / 400.0
    m_bad = dict(m, wall_flux_mismatch_rel=0.01, energy_residual_rel=0.5, temperature_bound_violations=2)
    fails = coupled.acceptance_failures(types.SimpleNamespace(converged=False, linear_residuals_pass=True), m_bad)
    assert fails == ["Newton iteration converged", "wall-flux mismatch <= 0.5 %", "global energy residual <= 1e-3",
                     "temperature within the wall temperatures"]
    assert coupled.acceptance_failures(result, dict(m, wall_flux_mismatch_rel=0.0, energy_residual_rel=0.0)) == []


@pytest.mark.parametrize("shape", [(13, 9, 8), (5, 7, 3), (16, 16, 16), (3, 20, 1)])
def test_coarse_source_blocks(shape):
    # blocks of np.array_split along each axis, each block summed by brute force
    from porerad import coupled

    a = np.random.default_rng(1).normal(size=shape)
    parts = [np.array_split(np.arange(n), min(8, n)) for n in shape]
    want = np.zeros(tuple(len(p) for p in parts))
    for i, xs in enumerate(parts[0]):
        for j, ys in enumerate(parts[1]):
            for k, zs in enumerate(parts[2]):
                want[i, j, k] = sum(a[x, y, z] for x in xs for y in ys for z in zs)
    got = coupled.coarse_source(a)
    assert got.shape == want.shape
    assert np.max(np.abs(got - want)) <= 1e-13 * np.max(np.abs(want))
