"""Cell-centred finite-volume heat conduction on a digital microstructure.

Drop-in for the reference module porerad/conduction.py (lines 24-120).  Each
cell carries its phase conductivity; neighbours exchange heat through the
harmonic-mean conductance 2 k1 k2 / (k1 + k2) * dx^2 / dx, y and z are
periodic, and a cell next to a wall couples to the wall temperature through
the half-cell conductance 2 k dx.  The steady balance is A T = b(T_h, T_c) + Q.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
from scipy.sparse import coo_matrix, csr_matrix


@dataclass
class ConductionSystem:
    """Conduction matrix A and the wall responses of an x-oriented grid (conduction.py:24-50)."""

    matrix: csr_matrix
    hot_rhs: np.ndarray
    cold_rhs: np.ndarray
    conductivity: np.ndarray
    H_m: float

    @property
    def shape(self) -> tuple[int, int, int]:
        return tuple(int(s) for s in self.conductivity.shape)

    def boundary_rhs(self, T_hot_K: float, T_cold_K: float) -> np.ndarray:
        """b contributed by the two fixed wall temperatures."""

        return self.hot_rhs * float(T_hot_K) + self.cold_rhs * float(T_cold_K)

    def wall_fluxes(self, temperature_K: np.ndarray, T_hot_K: float, T_cold_K: float) -> tuple[float, float]:
        """Conductive flux (W m^-2) into the slab at the hot wall and out at the cold wall."""

        k = self.conductivity
        T = np.asarray(temperature_K, dtype=np.float64).reshape(self.shape)
        dx = self.H_m / k.shape[0]
        q_hot = np.mean(2.0 * k[0] * (float(T_hot_K) - T[0]) / dx)
        q_cold = np.mean(2.0 * k[-1] * (T[-1] - float(T_cold_K)) / dx)
        return float(q_hot), float(q_cold)


def _harmonic(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    return 2.0 * a * b / (a + b + 1.0e-30)


def assemble_conduction(conductivity_x: np.ndarray, H_m: float) -> ConductionSystem:
    """A and the wall responses for an x-oriented conductivity field (conduction.py:57-97)."""

    k = np.ascontiguousarray(conductivity_x, dtype=np.float64)
    nx, ny, nz = (int(s) for s in k.shape)
    dx = float(H_m) / nx
    face_area = dx * dx
    index = np.arange(nx * ny * nz, dtype=np.int32).reshape(nx, ny, nz)
    diag = np.zeros(index.size, dtype=np.float64)
    rows, cols, vals = [], [], []

    def couple(lo_idx, hi_idx, lo_k, hi_k) -> None:
        lo, hi = lo_idx.reshape(-1), hi_idx.reshape(-1)
        g = np.ascontiguousarray(_harmonic(lo_k, hi_k) * face_area / dx).reshape(-1)
        diag[lo] += g
        diag[hi] += g
        rows.extend((lo, hi))
        cols.extend((hi, lo))
        vals.extend((-g, -g))

    couple(index[:-1], index[1:], k[:-1], k[1:])
    couple(index[:, :-1], index[:, 1:], k[:, :-1], k[:, 1:])
    couple(index[:, -1], index[:, 0], k[:, -1], k[:, 0])
    couple(index[:, :, :-1], index[:, :, 1:], k[:, :, :-1], k[:, :, 1:])
    couple(index[:, :, -1], index[:, :, 0], k[:, :, -1], k[:, :, 0])

    wall_scale = dx * dx / dx
    hot = np.zeros(index.size, dtype=np.float64)
    cold = np.zeros_like(hot)
    hot.reshape(k.shape)[0] = 2.0 * k[0] * wall_scale
    cold.reshape(k.shape)[-1] = 2.0 * k[-1] * wall_scale
    diag += hot + cold
    rows.append(index.reshape(-1))
    cols.append(index.reshape(-1))
    vals.append(diag)
    A = coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                   shape=(index.size, index.size)).tocsc().tocsr()
    A.sum_duplicates()
    A.sort_indices()
    return ConductionSystem(A, hot, cold, k, float(H_m))


def solve_conduction(system: ConductionSystem, T_hot_K: float, T_cold_K: float, *, rtol: float = 1.0e-11) -> np.ndarray:
    """Pure-conduction field by Jacobi-preconditioned CG from the linear profile (conduction.py:100-120).

    Runs on the device (``prh_coupled_cg``): ||b - A T||_2 <= rtol ||b||_2.
    """

    from . import _native

    rhs = system.boundary_rhs(T_hot_K, T_cold_K)
    nx, ny, nz = system.shape
    profile = float(T_hot_K) + (float(T_cold_K) - float(T_hot_K)) * ((np.arange(nx, dtype=np.float64) + 0.5) / nx)
    x0 = np.broadcast_to(profile[:, None, None], (nx, ny, nz)).reshape(-1).copy()
    max_iterations = max(50000, 20 * nx)
    D = _native.Coupled(system.matrix, system.hot_rhs, system.cold_rhs, 2)
    try:
        D.set(0, rhs)
        D.set(1, x0)
        iterations, relative = D.cg(0, 1, float(rtol), max_iterations)
        T = D.get(1)
    finally:
        D.close()
    if iterations >= max_iterations or not relative <= 1.0e-9:
        raise RuntimeError(f"conduction solve did not converge: iterations={iterations}, residual={relative:.3e}")
    return T.reshape(system.shape)
