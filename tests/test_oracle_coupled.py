"""The coupled-solve oracle against closed forms of the discretisation.

* Empty box: no internal surface, so radiation does not enter the cell
  balance; the conduction field is the exact linear profile, the conductive
  wall flux k dT / H and the radiative one the parallel-plate value
  sigma (T_h^4 - T_c^4) / (2/eps - 1).
* A plate spanning the cross section: by the periodic symmetry every cell
  column carries the same 1-D field, the plate's two faces exchange with
  the walls as parallel plates, and the discrete 1-D balance (harmonic-mean
  links, 2 k dx wall links, radiation into the two face cells) is solved
  here by an independent dense Newton iteration.  The coupled solve carries
  the 1e-8 radiosity tolerance of its Newton steps, so temperatures agree to
  ~1e-6 K and fluxes to ~1e-7 relative.
"""

import numpy as np
import pytest

from oracle import coupled_oracle as C
from oracle import exchange_oracle as X
from oracle import oracle as O

SIGMA = C.SIGMA


def _setup(phases, H, R=32, seed=5):
    P = O.build_surface_patches(phases, H)
    o, r, c = O.first_hit_counts(P, R, seed)
    op = X.OracleOperator(o, r, c, np.full(P.n_patches, R, np.int64), P.patch_area_m2)
    cond = C.Conduction(np.where(phases == 1, C.K_SOLID, C.K_GAS), H)
    return P, op, cond


def test_empty_box_is_conduction_plus_parallel_plates():
    phases = np.zeros((10, 4, 4), np.uint8)
    H, Th, Tc, eps = 0.01, 700.0, 300.0, 0.6
    P, op, cond = _setup(phases, H)
    T, ok, updates, final = C.solve_coupled(cond, op, P, Th, Tc, eps)
    assert ok and updates == 0
    nx = phases.shape[0]
    linear = Th + (Tc - Th) * (np.arange(nx) + 0.5) / nx
    assert np.allclose(T.reshape(phases.shape), linear[:, None, None], rtol=0, atol=1e-9)
    qch, qrh, qcc, qrc = C.wall_fluxes(cond, P, T, final, Th, Tc)
    assert qch == pytest.approx(C.K_GAS * (Th - Tc) / H, rel=1e-10)
    assert qcc == pytest.approx(C.K_GAS * (Th - Tc) / H, rel=1e-10)
    plates = SIGMA * (Th**4 - Tc**4) / (2.0 / eps - 1.0)
    assert qrh == pytest.approx(plates, rel=1e-10)
    assert qrc == pytest.approx(plates, rel=1e-10)


def _plate_1d(nx, lo, hi, H, Th, Tc, eps):
    """The 1-D discrete balance of one cell column (independent of the oracle)."""

    k = np.where((np.arange(nx) >= lo) & (np.arange(nx) < hi), C.K_SOLID, C.K_GAS)
    dx = H / nx
    g = 2.0 * k[:-1] * k[1:] / (k[:-1] + k[1:]) * dx  # face area dx^2 over dx
    gh, gc = 2.0 * k[0] * dx, 2.0 * k[-1] * dx
    f = SIGMA * dx * dx / (2.0 / eps - 1.0)
    T = np.linspace(Th, Tc, nx)

    def residual(T):
        r = np.zeros(nx)
        r[:-1] += g * (T[:-1] - T[1:])
        r[1:] += g * (T[1:] - T[:-1])
        r[0] += gh * (T[0] - Th)
        r[-1] += gc * (T[-1] - Tc)
        r[lo] -= f * (Th**4 - T[lo] ** 4)  # radiation into the plate's hot face
        r[hi - 1] -= f * (Tc**4 - T[hi - 1] ** 4)
        return r

    for _ in range(50):
        J = np.zeros((nx, nx))
        for i in range(nx - 1):
            J[i, i] += g[i]
            J[i, i + 1] -= g[i]
            J[i + 1, i + 1] += g[i]
            J[i + 1, i] -= g[i]
        J[0, 0] += gh
        J[-1, -1] += gc
        J[lo, lo] += 4.0 * f * T[lo] ** 3
        J[hi - 1, hi - 1] += 4.0 * f * T[hi - 1] ** 3
        step = np.linalg.solve(J, -residual(T))
        T = T + step
        if np.max(np.abs(step)) < 1e-13:
            break
    q_cond_hot = gh * (Th - T[0]) / (dx * dx)
    q_rad_hot = f * (Th**4 - T[lo] ** 4) / (dx * dx)
    return T, q_cond_hot, q_rad_hot


@pytest.mark.parametrize("eps", [0.3, 0.9])
def test_plate_matches_one_dimensional_balance(eps):
    nx, lo, hi = 12, 5, 7
    phases = np.zeros((nx, 4, 4), np.uint8)
    phases[lo:hi] = 1
    H, Th, Tc = 0.01, 700.0, 300.0
    P, op, cond = _setup(phases, H)
    T, ok, updates, final = C.solve_coupled(cond, op, P, Th, Tc, eps)
    assert ok and updates >= 1
    T1, qc1, qr1 = _plate_1d(nx, lo, hi, H, Th, Tc, eps)
    T3 = T.reshape(phases.shape)
    assert np.max(np.abs(T3 - T1[:, None, None])) < 1e-5
    qch, qrh, qcc, qrc = C.wall_fluxes(cond, P, T, final, Th, Tc)
    assert qch == pytest.approx(qc1, rel=1e-7)
    assert qrh == pytest.approx(qr1, rel=1e-7)
    # steady state: what enters at the hot wall leaves at the cold wall
    assert qch + qrh == pytest.approx(qcc + qrc, rel=1e-7)
