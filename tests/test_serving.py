"""Scheduling logic of the radiosity server, on the CPU with a stand-in engine.

The stand-in keeps the engine contract of prh_batch (64 slots,
admit_linear / step / channels / result / release, device bytes) and
finishes a slot after a fixed number of sweeps, with absorbed power = E / 4,
so admission order, continuous batching and HBM-budget eviction can be
checked exactly.  The solver numerics are covered by the GPU tests.
"""

import types

import numpy as np
import pytest

from porerad import serving


class StandInEngine:
    def __init__(self, geometry, epochs_needed=2, nbytes=100):
        self.geometry = geometry
        self.n = geometry.n
        self.bytes = nbytes
        self.busy = {}
        self.epochs_needed = epochs_needed
        self.closed = False
        self.log = []
        self.calls = []

    def admit_linear(self, T_hot, T_cold, eps, rtol):
        free = [q for q in range(serving.SLOTS) if q not in self.busy]
        if not free:
            raise ValueError("all slots are busy")
        q = free[0]
        E = self.geometry.emission(serving.RadiosityRequest(self.geometry.key, T_hot, T_cold, eps))
        self.busy[q] = [E, 0]
        self.log.append(("admit", q))
        return q

    def step(self, max_ticks=1):
        # up to max_ticks sweeps, returning early once a slot has finished (the engine's contract)
        self.calls.append(max_ticks)
        for _ in range(max_ticks):
            done = 0
            for q, rec in self.busy.items():
                rec[1] += 1
                if rec[1] >= self.epochs_needed:
                    done |= 1 << q
            if done:
                break
        return done, sum(1 << q for q in self.busy)

    def channels(self, q):
        E, ep = self.busy[q]
        net = 0.25 * E - E
        q_h, q_c, q_i = self.geometry.channels(net)
        return q_h, q_c, q_i, 0.0, ep

    def result(self, q):
        E, ep = self.busy.pop(q)
        return 0.25 * E, 0.0, ep

    def release(self, q):
        self.busy.pop(q)

    def close(self):
        self.closed = True


def _patches(nx=4, ny=3, nz=3):
    from oracle import oracle as O

    phases = np.zeros((nx, ny, nz), np.uint8)
    phases[1, 1, 1] = 1
    return O.build_surface_patches(phases, 0.004)


class Clock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        self.t += 0.001
        return self.t


def test_requests_finish_and_channels_follow_net_power():
    P = _patches()
    srv = serving.RadiosityServer(engine_factory=lambda g: StandInEngine(g), clock=Clock())
    g = srv.add_geometry("a", P, operator=None)
    reqs = [serving.RadiosityRequest("a", 700.0, 300.0, 0.6, arrival_s=0.0, patch_power=(i == 0)) for i in range(5)]
    replies = srv.run(reqs, real_time=False)
    assert len(replies) == 5 and srv.pending() == 0
    E = g.emission(reqs[0])
    net = 0.25 * E - E
    assert replies[0].q_h == pytest.approx(-np.sum(net[P.patch_class == 0]) / g.cross_section)
    assert np.array_equal(replies[0].net_power_W, net)
    assert all(r.net_power_W is None for r in replies[1:])
    assert all(r.sweeps == 2 for r in replies)


def test_emission_matches_prescribed_field_arithmetic():
    P = _patches(6, 4, 5)
    srv = serving.RadiosityServer(engine_factory=lambda g: StandInEngine(g), clock=Clock())
    g = srv.add_geometry("a", P, operator=None)
    req = serving.RadiosityRequest("a", 900.0, 300.0, 0.3)
    T_m, dT = 600.0, 600.0
    nx = P.solid.shape[0]
    axial = (T_m + dT / 2.0) - (np.arange(nx, dtype=np.float64) + 0.5) * dT / nx
    T = np.empty(P.n_patches)
    T[P.patch_class == 0] = T_m + dT / 2.0
    T[P.patch_class == 1] = T_m - dT / 2.0
    inner = P.patch_class == 2
    T[inner] = np.broadcast_to(axial[:, None, None], P.solid.shape).ravel()[P.patch_cell[inner]]
    expect = 0.3 * serving.SIGMA * P.patch_area_m2 * ((T - T_m) * (T + T_m) * (T * T + T_m * T_m))
    assert np.array_equal(g.emission(req), expect)


def test_slots_are_capped_and_refilled_between_epochs():
    P = _patches()
    engines = []

    def factory(g):
        engines.append(StandInEngine(g, epochs_needed=3))
        return engines[-1]

    srv = serving.RadiosityServer(engine_factory=factory, clock=Clock())
    srv.add_geometry("a", P, operator=None)
    for _ in range(serving.SLOTS + 4):
        srv.submit(serving.RadiosityRequest("a", 700.0, 300.0, 0.6))
    srv.step()
    assert len(srv.slots["a"]) == serving.SLOTS and len(srv.queue) == 4
    srv.step()
    srv.step()  # the first SLOTS finish at their 3rd epoch
    assert len(srv.replies) == serving.SLOTS
    srv.step()  # the 4 waiting requests are admitted while nothing else runs
    assert len(srv.slots["a"]) == 4
    while srv.pending():
        srv.step()
    assert len(srv.replies) == serving.SLOTS + 4


def test_late_arrivals_join_running_batch():
    P = _patches()
    engines = []
    srv = serving.RadiosityServer(engine_factory=lambda g: engines.append(StandInEngine(g, 4)) or engines[-1],
                                  clock=Clock())
    srv.add_geometry("a", P, operator=None)
    srv.submit(serving.RadiosityRequest("a", 700.0, 300.0, 0.6))
    srv.step()
    srv.submit(serving.RadiosityRequest("a", 500.0, 300.0, 0.9))
    srv.step()
    assert len(srv.slots["a"]) == 2  # the late request runs beside the first one
    while srv.pending():
        srv.step()
    assert [r.request.T_hot_K for r in srv.replies] == [700.0, 500.0]


def test_hbm_budget_evicts_idle_engines_only():
    P = _patches()
    made = []

    def factory(g):
        made.append(StandInEngine(g, epochs_needed=2, nbytes=100))
        return made[-1]

    srv = serving.RadiosityServer(hbm_budget_bytes=150, engine_factory=factory, engine_bytes=lambda g: 100,
                                  other_bytes=lambda: 0,
                                  clock=Clock())
    srv.add_geometry("a", P, operator=None)
    srv.add_geometry("b", P, operator=None)
    srv.submit(serving.RadiosityRequest("a", 700.0, 300.0, 0.6))
    srv.submit(serving.RadiosityRequest("b", 700.0, 300.0, 0.6))
    srv.step()
    # only one engine fits: b waits while a runs
    assert set(srv.engines) == {"a"} and len(srv.queue) == 1
    while srv.pending():
        srv.step()
        assert srv.engine_bytes() <= 150
    assert len(srv.replies) == 2 and srv.evictions == 1 and made[0].closed


def test_unknown_geometry_is_rejected():
    srv = serving.RadiosityServer(engine_factory=lambda g: StandInEngine(g))
    with pytest.raises(KeyError):
        srv.submit(serving.RadiosityRequest("nope", 700.0, 300.0, 0.6))


def test_poisson_arrivals_have_the_requested_rate():
    t = serving.poisson_arrivals(200.0, 20000, 7)
    assert np.all(np.diff(t) > 0)
    assert t[-1] / 20000 == pytest.approx(1 / 200.0, rel=0.03)


def _coupled_runner(log, pause=0.005):
    import time

    def runner(g, req):
        log.append(req.T_hot_K)
        time.sleep(pause)
        return True, 2, {"q_total_hot_W_m2": 0.5 * req.T_hot_K}, ""

    return runner


def test_coupled_requests_run_beside_radiosity():
    P = _patches()
    log = []
    srv = serving.RadiosityServer(engine_factory=lambda g: StandInEngine(g, 3), clock=Clock(),
                                  coupled_runner=_coupled_runner(log), coupled_bytes=lambda g: 10)
    srv.add_geometry("a", P, operator=None)
    reqs = [serving.RadiosityRequest("a", 700.0, 300.0, 0.6) for _ in range(20)]
    reqs += [serving.CoupledRequest("a", 700.0, 300.0, 0.6), serving.CoupledRequest("a", 900.0, 300.0, 0.3)]
    replies = srv.run(reqs, real_time=False)
    srv.close()
    assert len(replies) == len(reqs) and srv.pending() == 0
    coupled = [r for r in replies if isinstance(r, serving.CoupledReply)]
    assert log == [700.0, 900.0]  # one at a time, in arrival order
    assert [c.metrics["q_total_hot_W_m2"] for c in coupled] == [350.0, 450.0]
    assert all(c.converged and c.newton_updates == 2 and c.latency_s > 0 for c in coupled)


def test_coupled_request_waits_for_hbm_budget():
    P = _patches()
    log = []
    srv = serving.RadiosityServer(hbm_budget_bytes=150, engine_factory=lambda g: StandInEngine(g, 3, nbytes=100),
                                  other_bytes=lambda: 0,
                                  engine_bytes=lambda g: 100, clock=Clock(), coupled_runner=_coupled_runner(log),
                                  coupled_bytes=lambda g: 100)
    srv.add_geometry("a", P, operator=None)
    srv.submit(serving.RadiosityRequest("a", 700.0, 300.0, 0.6))
    srv.submit(serving.CoupledRequest("a", 700.0, 300.0, 0.6))
    srv.step()
    # the engine holds 100 of 150 bytes while it runs: the coupled solve must wait
    assert log == [] and len(srv.coupled_queue) == 1
    while srv.pending():
        srv.step()
        assert srv.engine_bytes() <= 150
    srv.close()
    assert log == [700.0] and srv.evictions == 1


def test_coupled_runner_errors_reach_the_caller():
    P = _patches()

    def runner(g, req):
        raise RuntimeError("radiosity reflection series did not converge")

    srv = serving.RadiosityServer(engine_factory=lambda g: StandInEngine(g), clock=Clock(), coupled_runner=runner,
                                  coupled_bytes=lambda g: 1)
    srv.add_geometry("a", P, operator=None)
    with pytest.raises(RuntimeError, match="did not converge"):
        srv.run([serving.CoupledRequest("a", 700.0, 300.0, 0.6)], real_time=False)
    srv.close()


def test_memory_held_outside_the_server_counts_against_the_budget():
    # 2 engines of 100 B would fit 250 B, but 60 B held by the tracer's workspace leave room for one
    P = _patches()
    srv = serving.RadiosityServer(hbm_budget_bytes=250, engine_factory=lambda g: StandInEngine(g, 2, nbytes=100),
                                  engine_bytes=lambda g: 100, other_bytes=lambda: 60, clock=Clock())
    srv.add_geometry("a", P, operator=None)
    srv.add_geometry("b", P, operator=None)
    srv.submit(serving.RadiosityRequest("a", 700.0, 300.0, 0.6))
    srv.submit(serving.RadiosityRequest("b", 700.0, 300.0, 0.6))
    srv.step()
    assert set(srv.engines) == {"a"} and len(srv.queue) == 1
    while srv.pending():
        srv.step()
    assert len(srv.replies) == 2


def test_sweeps_per_step_reaches_the_engine():
    P = _patches()
    engines = []

    def factory(g):
        engines.append(StandInEngine(g, epochs_needed=5))
        return engines[-1]

    srv = serving.RadiosityServer(engine_factory=factory, clock=Clock(), sweeps_per_step=3)
    srv.add_geometry("a", P, operator=None)
    replies = srv.run([serving.RadiosityRequest("a", 700.0, 300.0, 0.5)], real_time=False)
    assert len(replies) == 1 and set(engines[0].calls) == {3}
    assert srv.sweeps == 3 * len(engines[0].calls)
    with pytest.raises(ValueError):
        serving.RadiosityServer(sweeps_per_step=0)
