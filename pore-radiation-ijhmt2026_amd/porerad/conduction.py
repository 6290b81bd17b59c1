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
from scipy.sparse import csr_matrix


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
        """Conductive flux (W m^-2) into the slab at the hot wall and out at the cold wall: the mean over
        the wall of the half-cell flux 2 k (T_wall - T_cell) / dx."""

        T = np.asarray(temperature_K, dtype=np.float64).reshape(self.shape)
        half_cell = self.H_m / self.shape[0]
        first, last = (self.conductivity[0], T[0]), (self.conductivity[-1], T[-1])
        into_hot = 2.0 * first[0] * (float(T_hot_K) - first[1]) / half_cell
        out_cold = 2.0 * last[0] * (last[1] - float(T_cold_K)) / half_cell
        return float(np.mean(into_hot)), float(np.mean(out_cold))


def _harmonic(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    return 2.0 * a * b / (a + b + 1.0e-30)


def assemble_conduction(conductivity_x: np.ndarray, H_m: float) -> ConductionSystem:
    """A and the wall responses for an x-oriented conductivity field (conduction.py:57-97).

    Built as a 7-point stencil table: every cell's face conductances to its
    six neighbours (harmonic means times dx^2 / dx; y and z periodic, none
    across the x ends), the diagonal as their sum plus the wall links
    2 k dx^2 / dx of the first and last layers, then each row's entries sorted
    by column straight into CSR form.  The diagonal adds the terms in the
    order the reference's assembly does (+x, -x, the two y faces, the two z
    faces, the walls), so the matrix is the reference's bit for bit; on grids
    with fewer than three cells across y or z, where two faces join the same
    pair of cells, their entries are summed in column order.
    """

    k = np.ascontiguousarray(conductivity_x, dtype=np.float64)
    nx, ny, nz = (int(v) for v in k.shape)
    cells = nx * ny * nz
    dx = float(H_m) / nx
    face_area = dx * dx
    # conductance of the face towards the next cell along each axis (x: not at the last layer)
    up_x = np.zeros_like(k)
    up_x[:-1] = _harmonic(k[:-1], k[1:]) * face_area / dx
    up_y = _harmonic(k, np.roll(k, -1, axis=1)) * face_area / dx
    up_z = _harmonic(k, np.roll(k, -1, axis=2)) * face_area / dx
    down_x = np.zeros_like(k)
    down_x[1:] = up_x[:-1]
    down_y = np.roll(up_y, 1, axis=1)
    down_z = np.roll(up_z, 1, axis=2)

    # diagonal: +x, -x, then the y faces and the z faces, each pair in the
    # reference's order (the last layer along an axis takes its wrap face last)
    last_y = (np.arange(ny) == ny - 1)[None, :, None]
    last_z = (np.arange(nz) == nz - 1)[None, None, :]
    diag = np.zeros_like(k)
    for term in (up_x, down_x, np.where(last_y, down_y, up_y), np.where(last_y, up_y, down_y),
                 np.where(last_z, down_z, up_z), np.where(last_z, up_z, down_z)):
        diag += term
    wall_link = 2.0 * k * (dx * dx / dx)
    hot = np.zeros_like(k)
    cold = np.zeros_like(k)
    hot[0] = wall_link[0]
    cold[-1] = wall_link[-1]
    diag += hot + cold

    # the 7 candidate entries of every row; missing x neighbours get a column past the end
    idx = np.arange(cells, dtype=np.int64).reshape(k.shape)
    past = np.int64(cells)
    cols = np.stack([
        np.where(np.arange(nx)[:, None, None] > 0, idx - ny * nz, past),
        np.where(np.arange(nx)[:, None, None] < nx - 1, idx + ny * nz, past),
        np.roll(idx, 1, axis=1), np.roll(idx, -1, axis=1),
        np.roll(idx, 1, axis=2), np.roll(idx, -1, axis=2),
        idx,
    ], axis=-1).reshape(cells, 7)
    vals = np.stack([-down_x, -up_x, -down_y, -up_y, -down_z, -up_z, diag], axis=-1).reshape(cells, 7)
    order = np.argsort(cols, axis=1, kind="stable")
    cols = np.take_along_axis(cols, order, axis=1).ravel()
    vals = np.take_along_axis(vals, order, axis=1).ravel()
    row = np.repeat(np.arange(cells, dtype=np.int64), 7)
    keep = cols < past
    cols, vals, row = cols[keep], vals[keep], row[keep]
    if min(ny, nz) < 3:
        # faces that join the same pair of cells share one entry
        first = np.ones(cols.size, dtype=bool)
        first[1:] = (cols[1:] != cols[:-1]) | (row[1:] != row[:-1])
        starts = np.flatnonzero(first)
        vals = np.add.reduceat(vals, starts)
        cols, row = cols[starts], row[starts]
    indptr = np.zeros(cells + 1, dtype=np.int64)
    np.cumsum(np.bincount(row, minlength=cells), out=indptr[1:])
    A = csr_matrix((vals, cols.astype(np.int32), indptr), shape=(cells, cells))
    A.has_sorted_indices = True
    return ConductionSystem(A, hot.reshape(-1), cold.reshape(-1), k, float(H_m))


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
