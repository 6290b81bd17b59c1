"""Projected view-factor operator and gray diffuse radiosity on MI355X.

Drop-in for the reference module porerad/exchange.py.  With raw view factors
F_ij = n_ij / R_i and patch areas A the exchange matrix is replaced by its
symmetric part plus a small diagonal support term,

    S = 0.5 (diag(A) F + F^T diag(A)) + delta diag(A),    delta = 1e-15,

and one positive diagonal factor d restores row closure, d * (S d) = A.  The
projected view factors F_p = diag(1/A) diag(d) S diag(d) are applied as
F_p x = (d / A) S(d x) and F_p^T x = d S(d x / A).

S lives in HBM as an FP64 CSR built by the HIP library (``prh_operator_*``);
every product with it, and the reflection series J = E + (1 - eps) F_p^T J,
runs on the GPU.  The balancing factor is found on the device as well
(``prh_operator_balance``): SciPy's Anderson mixing on y = log d restated
with device vectors (exchange.py:175-183), with the damped multiplicative
iteration as fallback.  The reciprocity check runs on the device too.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import _native
from .constants import SIGMA

SUPPORT_DIAGONAL_FRACTION = 1.0e-15
CLOSURE_TOLERANCE = 1.0e-10
_TINY = np.finfo(np.float64).tiny


def _merge_edges(origin, receiver, count, n: int):
    """Sort by (origin, receiver) and sum duplicate pairs (exchange.py:59-67)."""

    o = np.asarray(origin, dtype=np.int64)
    r = np.asarray(receiver, dtype=np.int64)
    c = np.asarray(count, dtype=np.int64)
    key = o * n + r
    if key.size and not np.all(key[1:] > key[:-1]):
        order = np.argsort(key, kind="stable")
        key, c = key[order], c[order]
        first = np.concatenate(([True], key[1:] != key[:-1]))
        starts = np.flatnonzero(first)
        c = np.add.reduceat(c, starts)
        key = key[starts]
    return (key // n).astype(np.int32), (key % n).astype(np.int32), c


class ExchangeOperator:
    """Projected view-factor matrix F_p built from integer first-hit counts (exchange.py:37-197).

    ``origin``, ``receiver`` and ``count`` hold the counts n_ij and
    ``rays_per_origin`` the rays R_i; every row must account for all of its
    rays (ValueError otherwise).  The projected rows must close to 1e-10
    (RuntimeError otherwise).
    """

    def __init__(
        self,
        origin: np.ndarray,
        receiver: np.ndarray,
        count: np.ndarray,
        rays_per_origin: np.ndarray,
        patch_area_m2: np.ndarray,
        *,
        support_diagonal_fraction: float = SUPPORT_DIAGONAL_FRACTION,
        balance_tolerance: float = 1.0e-11,
        event_chunk_size: int = 4_000_000,
    ) -> None:
        rays = np.asarray(rays_per_origin, dtype=np.int64)
        n = int(rays.size)
        self._origin, self._receiver, merged = _merge_edges(origin, receiver, count, n)
        observed = np.bincount(self._origin, weights=merged, minlength=n) if merged.size else np.zeros(n)
        if observed.shape[0] != n or not np.array_equal(observed.astype(np.int64), rays):
            raise ValueError("first-hit counts do not account for every emitted ray")
        self._count = merged
        self._weight = None
        self._setup(rays, patch_area_m2, support_diagonal_fraction, event_chunk_size)
        self.device = _native.Operator(n, self._origin, self._receiver, merged, rays, self.patch_area_m2,
                                       self.support_diagonal_fraction)
        self._finish(balance_tolerance)

    @classmethod
    def from_edges(
        cls,
        edges: _native.Edges,
        rays_per_origin: np.ndarray,
        patch_area_m2: np.ndarray,
        *,
        support_diagonal_fraction: float = SUPPORT_DIAGONAL_FRACTION,
        balance_tolerance: float = 1.0e-11,
        event_chunk_size: int = 4_000_000,
    ) -> "ExchangeOperator":
        """The same operator from a device-resident edge table (``view_factors.first_hit_edges``).

        The assembly runs on the device (``prh_operator_create_device``): the
        counts never reach the host.  ``origin``, ``receiver``, ``count`` and
        ``weight`` are copied back only when first read.
        """

        self = cls.__new__(cls)
        rays = np.asarray(rays_per_origin, dtype=np.int64)
        self._origin = self._receiver = self._count = self._weight = None
        self._setup(rays, patch_area_m2, support_diagonal_fraction, event_chunk_size)
        self.device = _native.Operator.from_edges(self.n_patches, edges, rays, self.patch_area_m2,
                                                  self.support_diagonal_fraction)
        self._finish(balance_tolerance)
        return self

    def _setup(self, rays, patch_area_m2, support_diagonal_fraction, event_chunk_size) -> None:
        self.rays_per_origin = rays
        self.patch_area_m2 = np.asarray(patch_area_m2, dtype=np.float64).copy()
        self.support_diagonal_fraction = float(support_diagonal_fraction)
        self.event_chunk_size = int(event_chunk_size)
        self.n_patches = int(rays.size)

    def _finish(self, balance_tolerance: float) -> None:
        self.scale, self.balance_method, self.balance_iterations = self._balance(float(balance_tolerance))
        self.row_sum_error = float(np.max(np.abs(self.matvec(np.ones(self.n_patches)) - 1.0)))
        if self.row_sum_error > CLOSURE_TOLERANCE:
            raise RuntimeError(f"projected rows do not close: {self.row_sum_error:.3e}")

    # -- the count table (host copies made on first use) ----------------------------
    def _edges(self) -> None:
        if self._origin is None:
            self._origin, self._receiver, self._count = self.device.edges()

    @property
    def origin(self) -> np.ndarray:
        self._edges()
        return self._origin

    @property
    def receiver(self) -> np.ndarray:
        self._edges()
        return self._receiver

    @property
    def count(self) -> np.ndarray:
        self._edges()
        return self._count

    @property
    def weight(self) -> np.ndarray:
        """count / R[origin] (exchange.py:73)."""

        if self._weight is None:
            self._weight = self.count.astype(np.float64) / self.rays_per_origin[self.origin]
        return self._weight

    # -- products (all on the GPU) -------------------------------------------------
    def raw_matvec(self, x: np.ndarray) -> np.ndarray:
        """F x (exchange.py:82-93)."""

        return self.device.apply(_native.APPLY_RAW, x)

    def raw_rmatvec(self, x: np.ndarray) -> np.ndarray:
        """F^T x (exchange.py:95-106)."""

        return self.device.apply(_native.APPLY_RAW_T, x)

    def symmetric_matvec(self, x: np.ndarray) -> np.ndarray:
        """S x (exchange.py:108-112)."""

        return self.device.apply(_native.APPLY_SYMMETRIC, x)

    def matvec(self, x: np.ndarray) -> np.ndarray:
        """F_p x = (d / A) S(d x) (exchange.py:114-117)."""

        return self.device.apply(_native.APPLY_MATVEC, x)

    def rmatvec(self, x: np.ndarray) -> np.ndarray:
        """F_p^T x = d S(d x / A) (exchange.py:119-122)."""

        return self.device.apply(_native.APPLY_RMATVEC, x)

    def set_sweep_order(self, order: np.ndarray | None) -> None:
        """Order in which the batched engine sweeps this operator's rows
        (``SurfacePatches.sweep_order()``, or None for the stored order).
        Rows swept together by one XCD then gather from nearby patches and
        share L2 lines; the engine's results change only in the rounding of
        its norm sums."""

        self.device.set_row_order(order)

    def reciprocity_error(self) -> float:
        """Largest relative defect of A_i F_p,ij = A_j F_p,ji over the count table's pairs (exchange.py:124-140).

        Evaluated on the device (``prh_operator_reciprocity``) from the stored
        symmetric matrix and the current balancing factor.
        """

        return self.device.reciprocity()

    # -- balancing ------------------------------------------------------------------
    def _balance(self, tolerance: float) -> tuple[np.ndarray, str, int]:
        """Solve d * (S d) = A for d > 0 (exchange.py:153-197), on the device.

        ``prh_operator_balance`` runs the reference's steps: no balancing when
        d = 1 closes the rows, else SciPy's Anderson mixing (M=10, w0=0.01,
        maxiter=500, f_tol=tolerance, Armijo line search) on y = log d with
        the reference's acceptance test, else the damped multiplicative
        iteration; RuntimeError when none converges.  Vectors and products stay
        on the device; the factor is left set on the device operator.
        """

        method, iterations = self.device.balance(tolerance)
        return self.device.get_scale(), ("none", "anderson", "multiplicative")[method], iterations


@dataclass
class RadiosityResult:
    """Emitted, absorbed and net (absorbed minus emitted) patch power (exchange.py:200-214)."""

    emitted_power_W: np.ndarray
    absorbed_power_W: np.ndarray
    net_power_W: np.ndarray
    residual_power_W: float
    iterations: int

    @property
    def global_energy_residual_W(self) -> float:
        """Net power of all patches plus the power still in flight."""

        return float(np.sum(self.net_power_W) + self.residual_power_W)


def _emissivity_vector(emissivity, n: int) -> np.ndarray:
    return np.ascontiguousarray(np.broadcast_to(np.asarray(emissivity, dtype=np.float64), (n,)))


def reflection_series(
    operator: ExchangeOperator,
    emitted_W: np.ndarray,
    emissivity: np.ndarray,
    *,
    relative_tolerance: float,
    max_reflections: int = 10000,
) -> RadiosityResult:
    """Absorbed power after all diffuse reflections of a (possibly signed) emission (exchange.py:217-245).

    Power in flight is passed on by F_p^T; each patch absorbs eps and reflects
    1 - eps.  Stops when ||flight||_1 <= relative_tolerance ||E||_1; raises
    RuntimeError after ``max_reflections`` reflections.
    """

    emitted = np.asarray(emitted_W, dtype=np.float64)
    n = operator.n_patches
    eps = _emissivity_vector(emissivity, n)
    absorbed, residual, iterations = operator.device.radiosity(emitted.reshape(1, n), eps, relative_tolerance,
                                                               max_reflections)
    absorbed = absorbed[0]
    return RadiosityResult(emitted, absorbed, absorbed - emitted, float(residual[0]), int(iterations[0]))


def reflection_series_batch(
    operator: ExchangeOperator,
    emitted_W: np.ndarray,
    emissivity: np.ndarray,
    *,
    relative_tolerance: float,
    max_reflections: int = 10000,
    precision: str = "fp64",
) -> list[RadiosityResult]:
    """``reflection_series`` for k emissions (rows of ``emitted_W``) in one device call.

    ``emissivity`` is one vector for all rows or one row per emission.
    ``precision="fp64"`` sums the reflection series in FP64 exactly as
    ``reflection_series``, one emission at a time; ``"fp64_batched"`` sums
    the same FP64 series for up to 32 emissions per pass over the operator
    (each stops on its own tolerance; sums in another order, so equal to
    rounding); ``"mixed"`` solves the same radiosity balance on the
    batched engine (64 slots; Chebyshev sweeps with FP16 gathered
    directions and FP32 recursions, FP64 residual refinement) until
    ||E - (I - (1-eps) F_p^T) J||_1 <= relative_tolerance ||E||_1 (the
    iteration count is then the number of sweeps).
    """

    E = np.ascontiguousarray(np.atleast_2d(emitted_W), dtype=np.float64)
    n = operator.n_patches
    eps = np.asarray(emissivity, dtype=np.float64)
    eps = np.ascontiguousarray(np.broadcast_to(eps, E.shape) if eps.ndim == 2 else _emissivity_vector(eps, n))
    mode = {"fp64": _native.FP64, "mixed": _native.MIXED, "fp64_batched": _native.FP64_BATCHED}[precision]
    absorbed, residual, iterations = operator.device.radiosity(E, eps, relative_tolerance, max_reflections, mode)
    return [RadiosityResult(E[q], absorbed[q], absorbed[q] - E[q], float(residual[q]), int(iterations[q]))
            for q in range(E.shape[0])]


def emission(emissivity, area_m2, temperature_K, reference_temperature_K: float | None = None) -> np.ndarray:
    """eps sigma A T^4, or eps sigma A (T - T_ref)(T + T_ref)(T^2 + T_ref^2) (exchange.py:265-270)."""

    T = np.asarray(temperature_K, dtype=np.float64)
    if reference_temperature_K is None:
        return emissivity * SIGMA * area_m2 * T**4
    ref = float(reference_temperature_K)
    return emissivity * SIGMA * area_m2 * ((T - ref) * (T + ref) * (T * T + ref * ref))


def net_patch_power(
    operator: ExchangeOperator,
    temperature_K: np.ndarray,
    emissivity: np.ndarray,
    area_m2: np.ndarray,
    *,
    relative_tolerance: float,
    reference_temperature_K: float | None = None,
) -> RadiosityResult:
    """Net radiative power of every patch for given patch temperatures (exchange.py:248-271).

    With ``reference_temperature_K`` the isothermal state at that temperature
    (zero net power) is subtracted exactly before the series is summed.
    """

    emitted = emission(emissivity, area_m2, temperature_K, reference_temperature_K)
    return reflection_series(operator, emitted, emissivity, relative_tolerance=relative_tolerance)


def net_patch_power_jvp(
    operator: ExchangeOperator,
    temperature_K: np.ndarray,
    delta_temperature_K: np.ndarray,
    emissivity: np.ndarray,
    area_m2: np.ndarray,
    *,
    relative_tolerance: float,
) -> np.ndarray:
    """Derivative of the net patch power along a patch temperature perturbation (exchange.py:274-287)."""

    derivative = 4.0 * emissivity * SIGMA * area_m2 * temperature_K**3 * delta_temperature_K
    response = reflection_series(operator, derivative, emissivity, relative_tolerance=relative_tolerance)
    return response.absorbed_power_W - derivative
