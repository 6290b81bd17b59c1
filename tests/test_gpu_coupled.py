"""The device coupled radiation-conduction solve (prh_coupled_*) on the GPU.

* the pure-conduction field against the oracle's SciPy CG;
* a plate spanning the cross section against the independent 1-D balance
  of tests/test_oracle_coupled.py (same tolerances);
* a random medium against the CPU oracle's Newton-Krylov solve: the same
  number of Newton updates, temperatures within 1e-8 K, wall fluxes within
  1e-9 of the total flux;
* archived canonical N=128 rows of results/coupled.csv (4 ray blocks of 64
  rays per patch, K_SOLID 0.20): every flux within 1e-9 of the larger total
  flux, T_min/T_max within 1e-8 K, the same Newton update count and an
  accepted solution.
"""

from __future__ import annotations

import numpy as np
import pytest

from golden_data import cases, phases_x
from oracle import coupled_oracle as C
from oracle import exchange_oracle as X
from oracle import oracle as O
from test_oracle_coupled import _plate_1d

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mods():
    from porerad import _native, conduction, coupled, exchange, view_factors
    from porerad.constants import K_GAS, K_SOLID

    _native.lib()
    return view_factors, exchange, conduction, coupled, (K_SOLID, K_GAS)


def _product(mods, phases, H, R, seeds):
    vf, ex, cd, _, (ks, kg) = mods
    P = vf.build_surface_patches(phases, H)
    o, r, c = vf.first_hit_counts_pooled(P, R, seeds)
    op = ex.ExchangeOperator(o, r, c, np.full(P.n_patches, R * len(seeds), np.int64), P.patch_area_m2)
    system = cd.assemble_conduction(np.where(phases == 1, ks, kg), H)
    return P, op, system


def test_conduction_matches_oracle(mods):
    _, _, cd, _, (ks, kg) = mods
    rng = np.random.default_rng(3)
    phases = (rng.random((16, 10, 10)) < 0.3).astype(np.uint8)
    k = np.where(phases == 1, ks, kg)
    system = cd.assemble_conduction(k, 0.01)
    T = cd.solve_conduction(system, 700.0, 300.0).ravel()
    ref = C.Conduction(k, 0.01)
    assert (system.matrix != ref.A).nnz == 0
    b = ref.rhs(700.0, 300.0)
    assert np.linalg.norm(ref.A @ T - b) <= 1e-10 * np.linalg.norm(b)
    assert np.max(np.abs(T - ref.solve(700.0, 300.0))) <= 1e-6


@pytest.mark.parametrize("shape", [(12, 9, 7), (6, 3, 3), (5, 2, 6), (4, 6, 1)])
def test_conduction_product_bitwise(mods, shape):
    """A x on the device equals SciPy's CSR product bit for bit: (12, 9, 7) and (6, 3, 3) take the
    stencil form (each row summed in column order), the grids with a lateral extent below 3 the CSR rows."""

    from porerad import _native

    rng = np.random.default_rng(sum(shape))
    phases = (rng.random(shape) < 0.3).astype(np.uint8)
    P, op, system = _product(mods, phases, 0.01, 8, [5])
    D = _native.Coupled(system.matrix, system.hot_rhs, system.cold_rhs, 2, P.geometry(), op.device)
    x = rng.normal(size=system.matrix.shape[0]) * 300.0
    D.set(0, x)
    D.matvec(0, 1)
    assert np.array_equal(D.get(1), system.matrix @ x)


@pytest.mark.parametrize("eps", [0.3, 0.9])
def test_plate_matches_one_dimensional_balance(mods, eps):
    _, _, _, cp, _ = mods
    nx, lo, hi = 12, 5, 7
    phases = np.zeros((nx, 4, 4), np.uint8)
    phases[lo:hi] = 1
    H, Th, Tc = 0.01, 700.0, 300.0
    P, op, system = _product(mods, phases, H, 32, [5])
    T0 = mods[2].solve_conduction(system, Th, Tc)
    res = cp.solve_coupled(system, op, P, Th, Tc, eps, T0)
    assert res.converged and res.newton_iterations >= 1
    T1, qc1, qr1 = _plate_1d(nx, lo, hi, H, Th, Tc, eps)
    assert np.max(np.abs(res.temperature_K - T1[:, None, None])) < 1e-5
    m = cp.wall_flux_metrics(system, P, res, Th, Tc)
    assert m["q_cond_hot_W_m2"] == pytest.approx(qc1, rel=1e-7)
    assert m["q_rad_hot_W_m2"] == pytest.approx(qr1, rel=1e-7)
    assert not cp.acceptance_failures(res, m)


def test_random_medium_matches_oracle(mods):
    _, _, cd, cp, (ks, kg) = mods
    rng = np.random.default_rng(11)
    phases = (rng.random((14, 10, 10)) < 0.15).astype(np.uint8)
    H, Th, Tc, eps = 0.004, 900.0, 300.0, 0.6
    P, op, system = _product(mods, phases, H, 64, [2])
    T0 = cd.solve_conduction(system, Th, Tc)
    res = cp.solve_coupled(system, op, P, Th, Tc, eps, T0)
    m = cp.wall_flux_metrics(system, P, res, Th, Tc)
    # oracle: same counts (bit-identical tracer), CPU operator and solver
    Q = O.build_surface_patches(phases, H)
    o, r, c = O.first_hit_counts(Q, 64, 2)
    rop = X.OracleOperator(o, r, c, np.full(Q.n_patches, 64, np.int64), Q.patch_area_m2)
    cond = C.Conduction(np.where(phases == 1, C.K_SOLID, C.K_GAS), H)
    T, ok, updates, final = C.solve_coupled(cond, rop, Q, Th, Tc, eps)
    assert res.converged and ok and res.newton_iterations == updates
    assert np.max(np.abs(res.temperature_K.ravel() - T)) <= 1e-8
    qch, qrh, qcc, qrc = C.wall_fluxes(cond, Q, T, final, Th, Tc)
    scale = max(abs(qch + qrh), abs(qcc + qrc))
    for got, want in ((m["q_cond_hot_W_m2"], qch), (m["q_rad_hot_W_m2"], qrh), (m["q_cond_cold_W_m2"], qcc),
                      (m["q_rad_cold_W_m2"], qrc)):
        assert abs(got - want) <= 1e-9 * scale


ARCHIVED = ["coupled_FA_x_T400_eps0.6_N128", "coupled_OF_x_T700_eps0.3_N128", "coupled_TL_z_T700_eps0.6_N128"]


@pytest.mark.parametrize("case_id", ARCHIVED)
def test_archived_coupled_rows(mods, case_id):
    vf, ex, cd, cp, (ks, kg) = mods
    row = next(c for c in cases()["coupled_canonical"] if c["case_id"] == case_id)
    hc = next(h for h in cases()["hash_cases"]
              if h["microstructure"] == row["microstructure"] and h["axis"] == row["transfer_axis"])
    phases = phases_x(row["microstructure"], row["transfer_axis"])
    seeds = hc["seeds_A"] + hc["seeds_B"]
    P, op, system = _product(mods, phases, row["H_m"], hc["rays_per_patch_per_block"], seeds)
    Th = row["T_mean_K"] + row["delta_T_K"] / 2.0
    Tc = row["T_mean_K"] - row["delta_T_K"] / 2.0
    T0 = cd.solve_conduction(system, Th, Tc)
    res = cp.solve_coupled(system, op, P, Th, Tc, row["epsilon"], T0)
    m = cp.wall_flux_metrics(system, P, res, Th, Tc)
    assert not cp.acceptance_failures(res, m)
    assert res.newton_iterations == int(row["newton_updates"])
    scale = max(abs(row["q_total_hot_W_m2"]), abs(row["q_total_cold_W_m2"]))
    for key in ("q_cond_hot_W_m2", "q_rad_hot_W_m2", "q_total_hot_W_m2", "q_cond_cold_W_m2", "q_rad_cold_W_m2",
                "q_total_cold_W_m2"):
        assert abs(m[key] - row[key]) <= 1e-9 * scale, key
    assert abs(m["T_min_K"] - row["T_min_K"]) <= 1e-8
    assert abs(m["T_max_K"] - row["T_max_K"]) <= 1e-8


@pytest.mark.parametrize("eps", [0.3, 0.6])
def test_mixed_precision_jacobian_solves_the_same_equations(mods, eps):
    """FP32-sweep Jacobian-vector products (FP64 refinement to 1e-8) leave the
    Newton residual in FP64, so the field solves the same discrete equations:
    every wall flux within 1e-8 of the total flux of the FP64 solve and the
    temperatures within 1e-6 K."""

    _, _, cd, cp, _ = mods
    rng = np.random.default_rng(21)
    phases = (rng.random((18, 14, 14)) < 0.15).astype(np.uint8)
    H, Th, Tc = 0.004, 800.0, 300.0
    P, op, system = _product(mods, phases, H, 96, [4, 5])
    T0 = cd.solve_conduction(system, Th, Tc)
    ref = cp.solve_coupled(system, op, P, Th, Tc, eps, T0)
    mixed = cp.solve_coupled(system, op, P, Th, Tc, eps, T0, precision="mixed")
    assert ref.converged and mixed.converged
    assert abs(mixed.newton_iterations - ref.newton_iterations) <= 1
    m_ref = cp.wall_flux_metrics(system, P, ref, Th, Tc)
    m_mix = cp.wall_flux_metrics(system, P, mixed, Th, Tc)
    scale = max(abs(m_ref["q_total_hot_W_m2"]), abs(m_ref["q_total_cold_W_m2"]))
    for key in ("q_cond_hot_W_m2", "q_rad_hot_W_m2", "q_cond_cold_W_m2", "q_rad_cold_W_m2"):
        assert abs(m_mix[key] - m_ref[key]) <= 1e-8 * scale, key
    assert np.max(np.abs(mixed.temperature_K - ref.temperature_K)) <= 1e-6
    assert not cp.acceptance_failures(mixed, m_mix)
