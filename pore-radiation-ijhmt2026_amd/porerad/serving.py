"""Continuous-batching server of radiosity requests over device-resident exchange operators.

A request names a resident geometry (its ``SurfacePatches`` and
``ExchangeOperator``), the two wall temperatures and the emissivity; it asks
for the radiation channels of the prescribed linear field between the walls
(porerad/coupled.py:170-194), solved to a relative tolerance with the
batched mixed-precision engine (``prh_batch_*``: Chebyshev sweeps over 64
slots per operator with FP16 gathered directions, FP64 refinement).  The emission is formed and the
channels are summed on the device, so a request moves a few scalars over
PCIe; the patch powers come back only when a request asks for them.

Scheduling.  Every geometry has one engine with 64 slots.  Each scheduler
step (1) admits queued requests, oldest first, into free slots of their
geometry's engine, creating the engine first if its device memory fits the
HBM budget (idle engines are released to make room); (2) advances every
engine that has running slots by one sweep (the engines queue their sweeps
on one stream, each waiting only for its previous sweep, so the device
always has work queued); (3) retires finished slots and replies.  Requests
that arrive while others are on the GPU join at the next sweep, so their
work overlaps on the device with the requests already running; a solve's
bits do not depend on what else is in its batch.

Coupled requests (``CoupledRequest``: the strictly coupled conduction-
radiation solve of coupled.py:428-490 with given phase conductivities) run
one at a time on a worker thread; the device solve has its own HIP stream,
so its kernels run beside the radiosity sweeps.  A coupled solve starts only
when its device working set fits the HBM budget next to the engines.

The engine type and the coupled runner are injectable (``engine_factory``,
``coupled_runner``) so that the admission and retirement logic can be
exercised without a GPU.
"""

from __future__ import annotations

import collections
import dataclasses
import threading
import time
from typing import Callable

import numpy as np

from .constants import SIGMA
from .view_factors import COLD, HOT, INTERNAL

SLOTS = 64
DEV_SLOT_BYTES = 72  # device schedule state of one slot


@dataclasses.dataclass
class RadiosityRequest:
    """Channels of the linear field between T_hot and T_cold with uniform emissivity."""

    key: str
    T_hot_K: float
    T_cold_K: float
    epsilon: float
    relative_tolerance: float = 1.0e-9
    arrival_s: float = 0.0
    tag: object = None
    patch_power: bool = False  # also return the net power of every patch


@dataclasses.dataclass
class CoupledRequest:
    """Strictly coupled conduction-radiation solve between T_hot and T_cold (coupled.py:428-490)."""

    key: str
    T_hot_K: float
    T_cold_K: float
    epsilon: float
    conductivity_solid: float = 1.0
    conductivity_gas: float = 0.026
    arrival_s: float = 0.0
    tag: object = None
    precision: str = "fp64"  # radiosity of the Jacobian-vector products: "fp64" series or "mixed" engine
    preconditioner_steps: tuple = (10, 20)  # Jacobi-PCG step counts tried in turn (coupled.solve_coupled)


@dataclasses.dataclass
class CoupledReply:
    request: CoupledRequest
    converged: bool
    newton_updates: int
    metrics: dict
    failure: str
    admitted_s: float
    finished_s: float

    @property
    def latency_s(self) -> float:
        return self.finished_s - self.request.arrival_s


@dataclasses.dataclass
class RadiosityReply:
    request: RadiosityRequest
    q_h: float
    q_c: float
    q_i: float
    net_power_W: np.ndarray | None
    residual_power_W: float
    sweeps: int
    admitted_s: float
    finished_s: float

    @property
    def latency_s(self) -> float:
        return self.finished_s - self.request.arrival_s


class Geometry:
    """A resident geometry: patch table, operator and the cached emission ingredients."""

    def __init__(self, key: str, patches, operator) -> None:
        self.key = key
        self.patches = patches
        self.operator = operator
        nx, ny, nz = patches.solid.shape
        self.n = patches.n_patches
        cls = patches.patch_class
        self.hot = cls == HOT
        self.cold = cls == COLD
        self.internal = cls == INTERNAL
        # x index of the internal patches' owning cells, and i + 1/2
        self.nx = nx
        self.cell_x = np.zeros(self.n, np.int32)
        self.cell_x[self.internal] = patches.patch_cell[self.internal] // (ny * nz)
        self.half_index = np.zeros(self.n)
        self.half_index[self.internal] = self.cell_x[self.internal].astype(np.float64) + 0.5
        self.area = patches.patch_area_m2
        self.cross_section = float(ny * nz) * (patches.H_m / float(nx)) ** 2
        self.conduction: dict[tuple[float, float], object] = {}  # assembled systems by conductivities

    def emission(self, req: RadiosityRequest) -> np.ndarray:
        """eps sigma A (T^4 - T_m^4) of the prescribed linear field.

        Same operations as prescribed_field_channels: walls at T_m +/- dT/2,
        cells at (T_m + dT/2) - (i + 1/2) dT / nx (coupled.py:138-143,
        180-184), isothermal state at T_m subtracted (exchange.py:265-270).
        """

        T_m = 0.5 * (req.T_hot_K + req.T_cold_K)
        dT = req.T_hot_K - req.T_cold_K
        T = np.empty(self.n)
        T[self.hot] = T_m + dT / 2.0
        T[self.cold] = T_m - dT / 2.0
        T[self.internal] = (T_m + dT / 2.0) - self.half_index[self.internal] * dT / self.nx
        return req.epsilon * SIGMA * self.area * ((T - T_m) * (T + T_m) * (T * T + T_m * T_m))

    def channels(self, net: np.ndarray) -> tuple[float, float, float]:
        a = self.cross_section
        return (-float(np.sum(net[self.hot])) / a, -float(np.sum(net[self.cold])) / a,
                -float(np.sum(net[self.internal])) / a)


def device_engine(geometry: Geometry):
    """A 64-slot engine over the geometry's operator, sweeping its rows in the
    Morton order of the patches (rows an XCD sweeps together share L2 lines)."""

    from . import _native

    geometry.operator.set_sweep_order(geometry.patches.sweep_order())
    engine = _native.Batch(geometry.operator.device)
    engine.set_field(geometry.patches.patch_class, geometry.cell_x, geometry.nx, geometry.cross_section)
    return engine


def device_coupled_runner(geometry: Geometry, request: CoupledRequest) -> tuple[bool, int, dict, str]:
    """Pure-conduction start field, coupled solve and wall-flux metrics on the device."""

    from . import conduction, coupled

    P = geometry.patches
    key = (float(request.conductivity_solid), float(request.conductivity_gas))
    system = geometry.conduction.get(key)
    if system is None:
        k = np.where(np.asarray(P.solid) == 1, key[0], key[1])
        system = geometry.conduction[key] = conduction.assemble_conduction(k, P.H_m)
    Th, Tc = float(request.T_hot_K), float(request.T_cold_K)
    T0 = conduction.solve_conduction(system, Th, Tc)
    result = coupled.solve_coupled(system, geometry.operator, P, Th, Tc, float(request.epsilon), T0,
                                   precision=request.precision,
                                   preconditioner_steps=tuple(request.preconditioner_steps))
    metrics = coupled.wall_flux_metrics(system, P, result, Th, Tc)
    return bool(result.converged), int(result.newton_iterations), metrics, result.failure


def device_coupled_bytes(geometry: Geometry) -> int:
    """Device bytes of one coupled solve: 73 cell vectors, scratch and the 7-point conduction matrix,
    patch vectors, and the mixed-precision Jacobian solve's vectors (5 FP32 + 3 FP64 per patch)."""

    n_cells = int(np.prod(geometry.patches.solid.shape))
    return n_cells * (83 * 8 + 7 * 12 + 8) + geometry.n * (32 + 44) + (1 << 20)


def device_engine_bytes(geometry: Geometry) -> int:
    """Device bytes of one engine: 64 slots x (3 FP32 + 2 FP16 + 5 FP64 vectors) per patch, partial
    sums and totals, staging (2 FP64 per patch), field (5 B per patch), the slots' parameters and
    schedule states, and the sweep control word."""

    n = geometry.n
    blocks = ((n + 3) // 4 + 7) // 8 * 8
    return (n * SLOTS * (3 * 4 + 2 * 2 + 5 * 8) + max(SLOTS * blocks, 4 * ((n + 255) // 256)) * 8
            + 2 * SLOTS * 8 + n * 21 + SLOTS * 32 + SLOTS * DEV_SLOT_BYTES + 32)


def device_tracer_bytes() -> int:
    """Device bytes the tracer's workspace holds on the current device (traces run beside the server)."""

    from . import _native

    return _native.tracer_workspace()[0]


class RadiosityServer:
    """Scheduler of radiosity requests (see the module docstring)."""

    def __init__(self, hbm_budget_bytes: int | None = None,
                 engine_factory: Callable[[Geometry], object] = device_engine,
                 engine_bytes: Callable[[Geometry], int] = device_engine_bytes,
                 clock: Callable[[], float] = time.perf_counter,
                 coupled_runner: Callable[[Geometry, CoupledRequest], tuple] = device_coupled_runner,
                 coupled_bytes: Callable[[Geometry], int] = device_coupled_bytes,
                 other_bytes: Callable[[], int] = device_tracer_bytes,
                 sweeps_per_step: int = 1) -> None:
        """``other_bytes`` reports device memory held outside the server that counts against the budget
        (by default the tracer's workspace).  ``sweeps_per_step``: sweeps each busy engine runs per
        scheduler step (one replay of that many ticks); 1 admits and retires at every sweep, more
        saves host round trips at the cost of that granularity."""

        if sweeps_per_step < 1:
            raise ValueError("sweeps_per_step must be at least 1")
        self.sweeps_per_step = int(sweeps_per_step)
        self.budget = hbm_budget_bytes
        self.other_size = other_bytes
        self.engine_factory = engine_factory
        self.engine_size = engine_bytes
        self.clock = clock
        self.geometries: dict[str, Geometry] = {}
        self.engines: dict[str, object] = {}
        self.slots: dict[str, dict[int, tuple[RadiosityRequest, float]]] = {}
        self.queue: collections.deque[RadiosityRequest] = collections.deque()
        self.replies: list[RadiosityReply] = []
        self.sweeps = 0
        self.evictions = 0
        self.coupled_runner = coupled_runner
        self.coupled_size = coupled_bytes
        self.coupled_queue: collections.deque[CoupledRequest] = collections.deque()
        self._coupled_thread: threading.Thread | None = None
        self._coupled_job: tuple[CoupledRequest, float] | None = None
        self._coupled_done: list[CoupledReply] = []
        self._coupled_error: BaseException | None = None
        self._lock = threading.Lock()

    # -- registration ------------------------------------------------------------
    def add_geometry(self, key: str, patches, operator) -> Geometry:
        g = Geometry(key, patches, operator)
        self.geometries[key] = g
        self.slots[key] = {}
        return g

    def engine_stats(self) -> dict:
        """Tick timings of the live engines that report them (``Batch.stats``), by geometry."""

        out = {}
        for key, engine in self.engines.items():
            stats = getattr(engine, "stats", None)
            if stats is not None:
                out[key] = stats()
        return out

    def engine_bytes(self) -> int:
        """Device bytes of the live engines, of the running coupled solve and held outside the server."""

        total = sum(self.engine_size(self.geometries[k]) for k in self.engines) + int(self.other_size())
        job = self._coupled_job
        if job is not None:
            total += self.coupled_size(self.geometries[job[0].key])
        return total

    # -- admission -----------------------------------------------------------------
    def submit(self, request: RadiosityRequest | CoupledRequest) -> None:
        if request.key not in self.geometries:
            raise KeyError(f"unknown geometry {request.key!r}")
        if isinstance(request, CoupledRequest):
            self.coupled_queue.append(request)
        else:
            self.queue.append(request)

    # -- coupled solves --------------------------------------------------------------
    def _start_coupled(self) -> None:
        if self._coupled_job is not None or not self.coupled_queue:
            return
        req = self.coupled_queue[0]
        g = self.geometries[req.key]
        if self.budget is not None:
            need = self.coupled_size(g)
            while self.engine_bytes() + need > self.budget:
                idle = [k for k in self.engines if not self.slots[k]]
                if not idle:
                    return
                self.engines.pop(idle[0]).close()
                self.evictions += 1
        self.coupled_queue.popleft()
        self._coupled_job = (req, self.clock())
        self._coupled_thread = threading.Thread(target=self._run_coupled, args=(g, req), daemon=True)
        self._coupled_thread.start()

    def _run_coupled(self, g: Geometry, req: CoupledRequest) -> None:
        try:
            converged, updates, metrics, failure = self.coupled_runner(g, req)
            with self._lock:
                self._coupled_done.append(CoupledReply(req, converged, updates, metrics, failure,
                                                       self._coupled_job[1], self.clock()))
        except BaseException as exc:  # reported by the scheduler thread
            with self._lock:
                self._coupled_error = exc
        finally:
            with self._lock:
                self._coupled_job = None

    def _collect_coupled(self) -> list[CoupledReply]:
        with self._lock:
            done, self._coupled_done = self._coupled_done, []
            error, self._coupled_error = self._coupled_error, None
        if error is not None:
            raise error
        return done

    def _engine_for(self, key: str):
        if key in self.engines:
            return self.engines[key]
        need = self.engine_size(self.geometries[key])
        if self.budget is not None:
            while self.engine_bytes() + need > self.budget:
                idle = [k for k in self.engines if not self.slots[k]]
                if not idle:
                    return None
                self.engines.pop(idle[0]).close()
                self.evictions += 1
        engine = self.engine_factory(self.geometries[key])
        self.engines[key] = engine
        return engine

    def _admit(self) -> None:
        waiting = collections.deque()
        while self.queue:
            req = self.queue.popleft()
            busy = self.slots[req.key]
            if len(busy) >= SLOTS:
                waiting.append(req)
                continue
            engine = self._engine_for(req.key)
            if engine is None:
                waiting.append(req)
                continue
            slot = engine.admit_linear(req.T_hot_K, req.T_cold_K, req.epsilon, req.relative_tolerance)
            busy[slot] = (req, self.clock())
        self.queue = waiting

    # -- one scheduler step ----------------------------------------------------------
    def step(self) -> list:
        """One scheduler step; returns the replies (radiosity and coupled) finished in it."""

        self._admit()
        self._start_coupled()
        done_now: list = []
        for key, engine in list(self.engines.items()):
            busy = self.slots[key]
            if not busy:
                continue
            done, _ = engine.step(self.sweeps_per_step)
            self.sweeps += self.sweeps_per_step
            g = self.geometries[key]
            for slot in list(busy):
                if not (done >> slot) & 1:
                    continue
                req, admitted = busy.pop(slot)
                q_h, q_c, q_i, residual, sweeps = engine.channels(slot)
                net = None
                if req.patch_power:
                    absorbed, _, _ = engine.result(slot)
                    net = absorbed - g.emission(req)
                else:
                    engine.release(slot)
                done_now.append(RadiosityReply(req, q_h, q_c, q_i, net, residual, sweeps, admitted, self.clock()))
        done_now.extend(self._collect_coupled())
        if not done_now and not any(self.slots.values()) and self._coupled_job is not None:
            time.sleep(0.001)  # only a coupled solve runs: do not spin
        self.replies.extend(done_now)
        return done_now

    def pending(self) -> int:
        with self._lock:
            coupled = len(self._coupled_done) + (1 if self._coupled_job is not None else 0)
        return len(self.queue) + sum(len(b) for b in self.slots.values()) + len(self.coupled_queue) + coupled

    def run(self, requests: list[RadiosityRequest], *, real_time: bool = True) -> list[RadiosityReply]:
        """Serve requests whose ``arrival_s`` are offsets from now; returns the replies in completion order.

        With ``real_time`` a request is submitted only once its arrival time
        has passed on the clock (an open-loop arrival process); otherwise all
        are queued at once (closed loop, saturation throughput).
        """

        order = sorted(requests, key=lambda r: r.arrival_s)
        start = self.clock()
        first = len(self.replies)
        i = 0
        while i < len(order) or self.pending():
            now = self.clock() - start
            while i < len(order) and (not real_time or order[i].arrival_s <= now):
                req = order[i]
                req.arrival_s = start + (req.arrival_s if real_time else 0.0)
                self.submit(req)
                i += 1
            if self.pending():
                self.step()
            elif i < len(order):
                wait = order[i].arrival_s - (self.clock() - start)
                if wait > 0:
                    time.sleep(min(wait, 0.002))
        return self.replies[first:]

    def close(self) -> None:
        if self._coupled_thread is not None:
            self._coupled_thread.join()
        for e in self.engines.values():
            e.close()
        self.engines.clear()


def poisson_arrivals(rate_per_s: float, count: int, seed: int) -> np.ndarray:
    """Arrival offsets (s) of a Poisson process with the given rate."""

    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate_per_s, size=count))


def latency_summary(replies: list, wall_s: float) -> dict:
    lat = np.array([r.latency_s for r in replies])
    return {
        "solves": int(lat.size),
        "solves_per_s": float(lat.size / wall_s) if wall_s > 0 else float("nan"),
        "p50_ms": float(np.percentile(lat, 50) * 1e3),
        "p99_ms": float(np.percentile(lat, 99) * 1e3),
        "mean_ms": float(lat.mean() * 1e3),
    }
