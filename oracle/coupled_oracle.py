"""FP64 CPU oracle of the strictly coupled radiation-conduction solve -- TEST INFRASTRUCTURE ONLY.

A plain NumPy/SciPy restatement of porerad/conduction.py:24-120 (cell-centred
finite-volume conduction, harmonic-mean conductances, periodic y/z, 2 k dx
wall links, Jacobi-preconditioned CG) and porerad/coupled.py:45-126,
232-490 (patch/cell maps, radiation and its JVP, FGMRES with two modified
Gram-Schmidt passes, fixed Jacobi-PCG preconditioner, Newton-Krylov with a
bounded backtracking line search, wall-flux metrics), on top of the exchange
oracle (oracle/exchange_oracle.py).  Used only as the checker of the device
coupled solve on small grids.

Pinning: tests/test_oracle_coupled.py checks it against closed forms that
the discretisation fixes exactly -- an empty box (linear conduction profile,
parallel-plate radiation between the walls) and a plate spanning the cross
section (the 1-D conduction-radiation balance solved independently by a
scalar Newton iteration); the device path is pinned at N=128 against the
archived rows of results/coupled.csv (tests/test_gpu_coupled.py).
"""

from __future__ import annotations

import numpy as np
from scipy.sparse import coo_matrix
from scipy.sparse.linalg import LinearOperator, cg

from . import exchange_oracle as X

SIGMA = X.SIGMA
_TINY = np.finfo(np.float64).tiny
K_SOLID = 0.20
K_GAS = 0.026


# -- conduction (conduction.py:24-120) ------------------------------------------------

class Conduction:
    def __init__(self, conductivity_x, H_m):
        k = np.ascontiguousarray(conductivity_x, dtype=np.float64)
        nx, ny, nz = k.shape
        dx = float(H_m) / nx
        face = dx * dx
        index = np.arange(k.size).reshape(k.shape)
        diag = np.zeros(k.size)
        rows, cols, vals = [], [], []

        def couple(lo, hi, klo, khi):
            lo, hi = lo.reshape(-1), hi.reshape(-1)
            g = (2.0 * klo * khi / (klo + khi + 1.0e-30) * face / dx).reshape(-1)
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
        self.hot = np.zeros(k.size)
        self.cold = np.zeros(k.size)
        self.hot.reshape(k.shape)[0] = 2.0 * k[0] * (dx * dx / dx)
        self.cold.reshape(k.shape)[-1] = 2.0 * k[-1] * (dx * dx / dx)
        diag += self.hot + self.cold
        rows.append(index.reshape(-1))
        cols.append(index.reshape(-1))
        vals.append(diag)
        A = coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                       shape=(k.size, k.size)).tocsr()
        A.sum_duplicates()
        A.sort_indices()
        self.A, self.k, self.H, self.shape = A, k, float(H_m), k.shape

    def rhs(self, Th, Tc):
        return self.hot * Th + self.cold * Tc

    def solve(self, Th, Tc, rtol=1.0e-11):
        nx = self.shape[0]
        profile = Th + (Tc - Th) * ((np.arange(nx) + 0.5) / nx)
        x0 = np.broadcast_to(profile[:, None, None], self.shape).reshape(-1).copy()
        inv = 1.0 / self.A.diagonal()
        M = LinearOperator(self.A.shape, matvec=lambda v: inv * v, dtype=np.float64)
        T, info = cg(self.A, self.rhs(Th, Tc), x0=x0, rtol=rtol, atol=0.0, maxiter=max(50000, 20 * nx), M=M)
        assert info == 0
        return T

    def wall_fluxes(self, T, Th, Tc):
        k = self.k
        T = T.reshape(self.shape)
        dx = self.H / self.shape[0]
        return float(np.mean(2.0 * k[0] * (Th - T[0]) / dx)), float(np.mean(2.0 * k[-1] * (T[-1] - Tc) / dx))


# -- radiation (coupled.py:45-126) ------------------------------------------------------

def patch_temperature(patches, T_cells, Th, Tc):
    cls = patches.patch_class
    T = np.empty(patches.n_patches)
    T[cls == 0] = Th
    T[cls == 1] = Tc
    inner = cls == 2
    T[inner] = np.asarray(T_cells).ravel()[patches.patch_cell[inner]]
    return T


def cell_power(patches, patch_power):
    inner = patches.patch_class == 2
    total = np.zeros(patches.solid.size)
    np.add.at(total, patches.patch_cell[inner], patch_power[inner])
    return total


def emission_diagonal(patches, T_cells, eps):
    inner = patches.patch_class == 2
    owner = patches.patch_cell[inner]
    T = np.asarray(T_cells).ravel()[owner]
    d = np.zeros(patches.solid.size)
    np.add.at(d, owner, 4.0 * eps * SIGMA * patches.patch_area_m2[inner] * T**3)
    return d


def radiation(op, patches, T_cells, Th, Tc, eps, rtol):
    """(patch temperatures, emitted, absorbed, series residual, cell net power)."""

    Tp = patch_temperature(patches, T_cells, Th, Tc)
    e = np.full(patches.n_patches, float(eps))
    E = X.emission(e, patches.patch_area_m2, Tp)
    absorbed, residual, _ = X.reflection_series(op, E, e, relative_tolerance=rtol)
    return Tp, E, absorbed, residual, cell_power(patches, absorbed - E)


def radiation_jvp(op, patches, Tp, dT_cells, eps, rtol):
    inner = patches.patch_class == 2
    dTp = np.zeros(patches.n_patches)
    dTp[inner] = np.asarray(dT_cells).ravel()[patches.patch_cell[inner]]
    e = np.full(patches.n_patches, float(eps))
    dE = 4.0 * e * SIGMA * patches.patch_area_m2 * Tp**3 * dTp
    absorbed, _, _ = X.reflection_series(op, dE, e, relative_tolerance=rtol)
    return cell_power(patches, absorbed - dE)


# -- solvers (coupled.py:232-447) ------------------------------------------------------

def fgmres(matvec, b, precond, restart=30, max_iterations=90, rtol=1.0e-8, atol=1.0e-13):
    n = b.size
    x = np.zeros(n)
    b_norm = max(float(np.linalg.norm(b)), _TINY)
    goal = max(atol, rtol * b_norm)
    eps = np.finfo(np.float64).eps
    done = 0
    while done < max_iterations:
        r = b - matvec(x)
        beta = float(np.linalg.norm(r))
        if beta <= goal:
            return x, True, done
        m = min(restart, max_iterations - done)
        V, Z = np.zeros((m + 1, n)), np.zeros((m, n))
        H, g = np.zeros((m + 1, m)), np.zeros(m + 1)
        V[0], g[0] = r / beta, beta
        again = None
        for j in range(m):
            Z[j] = precond(V[j])
            w = matvec(Z[j])
            for _ in range(2):
                for i in range(j + 1):
                    h = float(np.dot(V[i], w))
                    H[i, j] += h
                    w -= h * V[i]
            H[j + 1, j] = h_next = float(np.linalg.norm(w))
            if h_next > eps * beta:
                V[j + 1] = w / h_next
            y = np.linalg.lstsq(H[: j + 2, : j + 1], g[: j + 2], rcond=None)[0]
            estimate = float(np.linalg.norm(g[: j + 2] - H[: j + 2, : j + 1] @ y))
            done += 1
            if estimate <= goal or h_next <= eps * beta:
                cand = x + y @ Z[: j + 1]
                if float(np.linalg.norm(b - matvec(cand))) <= goal:
                    return cand, True, done
                again = cand
                break
        if again is not None:
            x = again
            continue
        x = x + np.linalg.lstsq(H, g, rcond=None)[0] @ Z
    return x, float(np.linalg.norm(b - matvec(x))) <= goal, done


def jacobi_pcg(A, shift, steps):
    inv = 1.0 / (A.diagonal() + shift)

    def apply(rhs):
        x = np.zeros_like(rhs)
        r = rhs.copy()
        z = inv * r
        p = z.copy()
        rz = r @ z
        for step in range(steps):
            Ap = A @ p + shift * p
            alpha = rz / (p @ Ap)
            x += alpha * p
            r -= alpha * Ap
            z = inv * r
            if step + 1 < steps:
                rz_next = r @ z
                p = z + (rz_next / rz) * p
                rz = rz_next
        return x

    return apply


def newton_krylov(cond, op, patches, Th, Tc, eps, T0, steps):
    """Returns (T, converged, newton updates, final (Tp, E, absorbed, residual, Q))."""

    A, b = cond.A, cond.rhs(Th, Tc)
    pcg = jacobi_pcg(A, emission_diagonal(patches, T0, eps), steps)
    T = np.asarray(T0, np.float64).ravel().copy()
    state = radiation(op, patches, T, Th, Tc, eps, 1.0e-8)
    res = A @ T - b - state[4]
    last = 0.0 if max(float(np.max(np.abs(res))), _TINY) <= 1.0e-13 else np.inf
    updates = 0

    def finish(ok):
        return T, ok, updates, radiation(op, patches, T, Th, Tc, eps, 1.0e-13)

    for _ in range(8):
        r_norm = float(np.max(np.abs(res)))
        if r_norm <= 1.0e-13 and last <= 1.0e-10:
            return finish(True)
        Tp = state[0].copy()

        def J(v):
            return A @ v - radiation_jvp(op, patches, Tp, v, eps, 1.0e-8)

        step, ok, _ = fgmres(J, -res, pcg)
        if not ok:
            return finish(False)
        if r_norm <= 1.0e-13 and float(np.max(np.abs(step))) <= 1.0e-10:
            return finish(True)
        alpha, accepted = 1.0, False
        while alpha >= 2.0**-20:
            cand = T + alpha * step
            if np.all(np.isfinite(cand)) and cand.min() >= Tc and cand.max() <= Th:
                st = radiation(op, patches, cand, Th, Tc, eps, 1.0e-8)
                tr = A @ cand - b - st[4]
                if float(np.max(np.abs(tr))) <= (1.0 - 1.0e-4 * alpha) * r_norm:
                    accepted = True
                    break
            alpha *= 0.5
        if not accepted:
            return finish(False)
        last = float(np.max(np.abs(cand - T)))
        T, state, res = cand, st, tr
        updates += 1
        if float(np.max(np.abs(res))) <= 1.0e-13 and last <= 1.0e-10:
            return finish(True)
    return finish(False)


def solve_coupled(cond, op, patches, Th, Tc, eps):
    T0 = cond.solve(Th, Tc)
    out = None
    for steps in (10, 20):
        out = newton_krylov(cond, op, patches, Th, Tc, eps, T0, steps)
        if out[1]:
            break
    return out


def wall_fluxes(cond, patches, T, final, Th, Tc):
    """(q_cond_hot, q_rad_hot, q_cond_cold, q_rad_cold) in W m^-2 (coupled.py:450-477)."""

    nx, ny, nz = patches.solid.shape
    area = float(ny * nz) * (patches.H_m / float(nx)) ** 2
    _, E, absorbed, _, _ = final
    net = absorbed - E
    qh, qc = cond.wall_fluxes(T, Th, Tc)
    return qh, -float(np.sum(net[patches.patch_class == 0])) / area, qc, float(
        np.sum(net[patches.patch_class == 1])) / area
