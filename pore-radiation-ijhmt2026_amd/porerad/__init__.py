"""Pore-scale radiation and conduction in digital microstructures of porous insulation.

MI355X implementation of the reference package's radiation path behind its
Python API: the first-hit view-factor tracer, the exchange operator and
radiosity solves, the coupled radiation-conduction solve, the case drivers,
the multi-bounce Monte Carlo, the microstructure generators and the
morphology descriptors, each running in hand-written HIP kernels for gfx950
(libporerad_hip.so, loaded by porerad._native; there is no CPU fallback).
"""

__version__ = "1.0.0"
