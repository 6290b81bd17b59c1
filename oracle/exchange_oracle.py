"""FP64 CPU oracle of the exchange operator, reflection series and prescribed-field
channels -- TEST INFRASTRUCTURE ONLY.

A plain restatement of porerad/exchange.py:37-287 and porerad/coupled.py:45-194
(prescribed field) used as the checker of the HIP radiosity path.  The edge
products follow exchange.py:82-106 (per-edge weight n_ij / R_i, products
summed in edge order, as np.bincount does) through the C loops of
libporerad_oracle.so; the balancing uses scipy.optimize.anderson with the
reference's settings (exchange.py:175-183).

Pinning: archived q_h/q_c/q_i of results/prescribed_field.csv and
results/fullpath_crosscheck.csv (tests/test_oracle_exchange.py), the
parallel-plate closed form of an empty box and the zero net power of an
isothermal field.
"""

from __future__ import annotations

import numpy as np
from scipy.optimize import NoConvergence, anderson

from . import oracle as _o

SIGMA = 5.670374419e-8
_TINY = np.finfo(np.float64).tiny
_LOG_MAX = float(np.log(np.finfo(np.float64).max))
_LOG_MIN = float(np.log(np.nextafter(0.0, 1.0)))
CLOSURE_TOLERANCE = 1.0e-10


class OracleOperator:
    """exchange.py:37-197 restated (FP64, CPU)."""

    def __init__(self, origin, receiver, count, rays_per_origin, area, *, delta=1.0e-15, balance_tolerance=1.0e-11):
        rays = np.asarray(rays_per_origin, dtype=np.int64)
        n = int(rays.size)
        pair = np.asarray(origin, np.int64) * n + np.asarray(receiver, np.int64)
        order = np.argsort(pair, kind="stable")
        pair = pair[order]
        start = np.r_[0, np.flatnonzero(pair[1:] != pair[:-1]) + 1]
        count = np.add.reduceat(np.asarray(count, np.int64)[order], start)
        self.origin = np.ascontiguousarray((pair[start] // n).astype(np.int32))
        self.receiver = np.ascontiguousarray((pair[start] % n).astype(np.int32))
        observed = np.zeros(n, np.int64)
        np.add.at(observed, self.origin, count)
        if not np.array_equal(observed, rays):
            raise ValueError("first-hit counts do not account for every emitted ray")
        self.count = count
        self.weight = np.ascontiguousarray(count.astype(np.float64) / rays[self.origin])
        self.area = np.asarray(area, np.float64).copy()
        self.delta = float(delta)
        self.n = n
        self.scale, self.balance_method, self.balance_iterations = self._balance(float(balance_tolerance))
        self.row_sum_error = float(np.max(np.abs(self.matvec(np.ones(n)) - 1.0)))
        if self.row_sum_error > CLOSURE_TOLERANCE:
            raise RuntimeError(f"projected rows do not close: {self.row_sum_error:.3e}")

    def raw_matvec(self, x):
        x = np.ascontiguousarray(x, np.float64)
        y = np.empty(self.n)
        _o.lib().orc_raw_matvec(_o._p(self.origin), _o._p(self.receiver), _o._p(self.weight), self.weight.size,
                                self.n, _o._p(x), _o._p(y))
        return y

    def raw_rmatvec(self, x):
        x = np.ascontiguousarray(x, np.float64)
        y = np.empty(self.n)
        _o.lib().orc_raw_rmatvec(_o._p(self.origin), _o._p(self.receiver), _o._p(self.weight), self.weight.size,
                                 self.n, _o._p(x), _o._p(y))
        return y

    def symmetric_matvec(self, x):
        A = self.area
        return 0.5 * (A * self.raw_matvec(x) + self.raw_rmatvec(A * x)) + self.delta * A * x

    def matvec(self, x):
        return (self.scale / self.area) * self.symmetric_matvec(self.scale * x)

    def rmatvec(self, x):
        return self.scale * self.symmetric_matvec(self.scale * (x / self.area))

    def reciprocity_error(self):
        """exchange.py:124-140 restated: max |A_i F_p,ij - A_j F_p,ji| / max |A_i F_p,ij| over the count pairs."""

        n, d, A = self.n, self.scale, self.area
        fwd = self.origin.astype(np.int64) * n + self.receiver
        rev = self.receiver.astype(np.int64) * n + self.origin
        at = np.minimum(np.searchsorted(fwd, rev), fwd.size - 1)
        w_rev = np.where(fwd[at] == rev, self.weight[at], 0.0)
        i, j = self.origin, self.receiver
        s_ij = 0.5 * (A[i] * self.weight + A[j] * w_rev)
        s_ji = 0.5 * (A[j] * w_rev + A[i] * self.weight)
        diag = i == j
        s_ij[diag] += self.delta * A[i[diag]]
        s_ji[diag] += self.delta * A[j[diag]]
        f_ij = A[i] * ((d[i] / A[i]) * s_ij * d[j])
        f_ji = A[j] * ((d[j] / A[j]) * s_ji * d[i])
        return float(np.max(np.abs(f_ij - f_ji)) / max(float(np.max(np.abs(f_ij))), _TINY))

    def _closure_error(self, y):
        if not (np.all(np.isfinite(y)) and np.max(y) < _LOG_MAX and np.min(y) > _LOG_MIN):
            return float("inf")
        d = np.exp(y)
        row = d * self.symmetric_matvec(d)
        if not np.all(np.isfinite(row)) or np.any(row <= 0.0):
            return float("inf")
        return float(np.max(np.abs(row / self.area - 1.0)))

    def _balance(self, tol):
        n, log_area = self.n, np.log(self.area)
        y0 = np.zeros(n)
        if self._closure_error(y0) <= tol:
            return np.exp(y0), "none", 0
        it = [0]

        def f(y):
            y = np.asarray(y, np.float64)
            if not (np.all(np.isfinite(y)) and np.max(y) < _LOG_MAX and np.min(y) > _LOG_MIN):
                return np.full(n, np.inf)
            p = self.symmetric_matvec(np.exp(y))
            if not np.all(np.isfinite(p)) or np.any(p <= 0.0):
                return np.full(n, np.inf)
            return y + np.log(p) - log_area

        def cb(_x, _f):
            it[0] += 1

        try:
            y = np.asarray(anderson(f, y0, M=10, w0=0.01, maxiter=500, f_tol=tol, callback=cb), np.float64)
            if self._closure_error(y) <= min(max(10.0 * tol, 10.0 * np.finfo(float).eps), CLOSURE_TOLERANCE):
                return np.exp(y), "anderson", it[0]
        except (NoConvergence, ValueError, FloatingPointError, np.linalg.LinAlgError):
            pass
        y = y0.copy()
        for k in range(1, 20_001):
            d = np.exp(y)
            row = d * self.symmetric_matvec(d)
            if not np.all(np.isfinite(row)) or np.any(row <= 0.0):
                break
            y = y + 0.5 * (log_area - np.log(row))
            e = self._closure_error(y)
            if not np.isfinite(e):
                break
            if e <= min(tol, CLOSURE_TOLERANCE):
                return np.exp(y), "multiplicative", k
        raise RuntimeError("diagonal balancing of the exchange matrix did not converge")


def reflection_series(op: OracleOperator, emitted, emissivity, *, relative_tolerance, max_reflections=10000):
    """exchange.py:217-245 restated: returns (absorbed, residual, iterations)."""

    emitted = np.asarray(emitted, np.float64)
    absorbed = np.zeros_like(emitted)
    flight = emitted.copy()
    initial = max(float(np.sum(np.abs(emitted))), _TINY)
    for k in range(1, int(max_reflections) + 1):
        incident = op.rmatvec(flight)
        absorbed += emissivity * incident
        flight = (1.0 - emissivity) * incident
        if float(np.sum(np.abs(flight))) <= float(relative_tolerance) * initial:
            return absorbed, float(np.sum(flight)), k
    raise RuntimeError("radiosity reflection series did not converge")


def emission(emissivity, area, T, T_ref=None):
    T = np.asarray(T, np.float64)
    if T_ref is None:
        return emissivity * SIGMA * area * T**4
    ref = float(T_ref)
    return emissivity * SIGMA * area * ((T - ref) * (T + ref) * (T * T + ref * ref))


def prescribed_channels(op: OracleOperator, patches, T_mean, delta_T, epsilon):
    """coupled.py:170-194 restated: (q_h, q_c, q_i, closure, absorbed, emitted, iterations)."""

    nx, ny, nz = patches.solid.shape
    cls = patches.patch_class
    axial = (T_mean + delta_T / 2.0) - (np.arange(nx, dtype=np.float64) + 0.5) * delta_T / nx
    T_cells = np.broadcast_to(axial[:, None, None], (nx, ny, nz)).ravel()
    T = np.empty(patches.n_patches)
    T[cls == 0] = T_mean + delta_T / 2.0
    T[cls == 1] = T_mean - delta_T / 2.0
    inner = cls == 2
    T[inner] = T_cells[patches.patch_cell[inner]]
    eps = np.full(patches.n_patches, float(epsilon))
    E = emission(eps, patches.patch_area_m2, T, T_mean)
    absorbed, residual, iters = reflection_series(op, E, eps, relative_tolerance=1.0e-13)
    net = absorbed - E
    area = float(ny * nz) * (patches.H_m / float(nx)) ** 2
    q_h = -float(np.sum(net[cls == 0])) / area
    q_c = -float(np.sum(net[cls == 1])) / area
    q_i = -float(np.sum(net[cls == 2])) / area
    closure = abs(q_h + q_c + q_i) / max(abs(q_h), abs(q_c), abs(q_i), 1.0e-30)
    return q_h, q_c, q_i, closure, absorbed, E, iters
