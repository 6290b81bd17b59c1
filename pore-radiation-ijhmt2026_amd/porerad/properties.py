"""Equivalent radiative properties identified from pore-scale ray statistics (porerad/properties.py).

Rays enter the slab at normal incidence through the face x = 0 at uniformly
distributed transverse positions and travel through the pore space of the
digital microstructure (periodic in the two transverse directions).  The
first path segment of every ray gives either an interception by the solid or
a right-censored exposure for a ray that crosses the slab unintercepted; the
directional extinction coefficient is the right-censored maximum-likelihood
estimate of the rate of the survival law exp(-beta l).  At a solid surface a
ray is absorbed with probability eps, otherwise it is reflected into a
Lambertian direction about the surface normal and followed further; the
albedo omega is the reflected fraction of all surface interactions and the
asymmetry factor g the mean cosine between incoming and reflected directions.
Lengths are in cells.  Restart rules "article" and "hit_point" as in the
reference (properties.py:13-22).

The ray kernel runs on the device (``prh_mc_properties``); the first path
lengths come back per ray so the exposure sums are NumPy's, as in the
reference.
"""

from __future__ import annotations

import math

import numpy as np

from . import _native

SEEDS = (123, 456)
RESTART_RULES = ("article", "hit_point")
N_RAYS = 200_000
MAX_BOUNCES = 1000


def solid_along_x(phases: np.ndarray, axis: str) -> np.ndarray:
    """The solid indicator with the transfer axis moved to the first index (properties.py:39-42)."""

    solid = np.asarray(phases) == 1
    return np.ascontiguousarray(np.moveaxis(solid, "xyz".index(axis), 0).astype(np.uint8))


def right_censored_mle(events: np.ndarray, censored: np.ndarray) -> tuple[float, float]:
    """Rate estimate events / total exposure and its standard error (properties.py:45-51)."""

    exposure = float(np.sum(events) + np.sum(censored))
    n_events = int(events.size)
    if exposure <= 0.0 or n_events == 0:
        return 0.0, 0.0
    return float(n_events / exposure), float(math.sqrt(n_events) / exposure)


def occupancy(phases: np.ndarray, axis: str) -> "_native.Geometry":
    """The device occupancy of the solid with the transfer axis first (all the ray kernels read)."""

    return _native.Geometry.occupancy_only(solid_along_x(phases, axis))


def trace_properties(phases: np.ndarray, axis: str, epsilon: float, *, n_rays: int = N_RAYS, seed: int = SEEDS[0],
                     max_bounces: int = MAX_BOUNCES, restart: str = "article", _occupancy=None) -> dict:
    """beta (per cell), omega and g for one geometry, transfer axis, emissivity and seed (properties.py:54-307)."""

    if restart not in RESTART_RULES:
        raise ValueError(f"restart must be one of {RESTART_RULES}")
    n = int(n_rays)
    events_all = np.empty(n, np.float64)
    censored_all = np.empty(n, np.float64)
    counters = np.zeros(2, np.int64)
    sums = np.zeros(2, np.float64)
    geo = _occupancy if _occupancy is not None else occupancy(phases, axis)
    try:
        _native.check(_native.lib().prh_mc_properties(
            geo.handle, n, int(seed) & 0xFFFFFFFF, float(epsilon), int(max_bounces),
            1 if restart == "hit_point" else 0, _native.ptr(events_all), _native.ptr(censored_all),
            _native.ptr(counters), _native.ptr(sums)), "prh_mc_properties")
    finally:
        if _occupancy is None:
            geo.close()
    events = events_all[events_all >= 0.0]
    censored = censored_all[censored_all >= 0.0]
    beta, beta_se = right_censored_mle(events, censored)
    interactions, scattered = int(counters[0]), int(counters[1])
    omega = scattered / interactions if interactions else 0.0
    omega_se = math.sqrt(omega * (1.0 - omega) / interactions) if interactions else 0.0
    g, g_se = 0.0, 0.0
    if scattered > 0:
        g = float(sums[0] / scattered)
        if scattered > 1:
            g_se = math.sqrt(max(0.0, float((sums[1] - scattered * g * g) / (scattered - 1))) / scattered)
    return {"n_rays": n, "beta": beta, "beta_se": beta_se, "omega": omega, "omega_se": omega_se, "g": g, "g_se": g_se}


def seed_average(values, ses):
    """Mean over seeds; the standard error combines within-seed and between-seed variance (properties.py:310-316)."""

    values, ses = np.asarray(values, float), np.asarray(ses, float)
    mean = float(values.mean())
    within = float(np.sum(ses**2) / values.size**2)
    between = float(values.var(ddof=1) / values.size) if values.size > 1 else 0.0
    return mean, float(np.sqrt(max(0.0, within + between)))


def identify_properties(phases: np.ndarray, axis: str, epsilon: float, *, seeds=SEEDS, n_rays: int = N_RAYS,
                        restart: str = "article") -> dict:
    """Properties of one geometry, axis and emissivity, averaged over independent seeds (properties.py:319-329)."""

    geo = occupancy(phases, axis)  # one upload serves every seed
    try:
        runs = [trace_properties(phases, axis, epsilon, n_rays=n_rays, seed=s, restart=restart, _occupancy=geo)
                for s in seeds]
    finally:
        geo.close()
    out = {"n_rays": int(sum(r["n_rays"] for r in runs))}
    for key in ("beta", "omega", "g"):
        mean, se = seed_average([r[key] for r in runs], [r[f"{key}_se"] for r in runs])
        out[key], out[f"{key}_se"] = mean, se
    out["beta_ci95_low"] = out["beta"] - 1.96 * out["beta_se"]
    out["beta_ci95_high"] = out["beta"] + 1.96 * out["beta_se"]
    return out
