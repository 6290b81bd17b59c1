"""Pore-scale radiation and conduction in digital microstructures of porous insulation.

MI355X implementation of the radiation hot path of the reference package:
the first-hit view-factor tracer and the radiosity exchange solve run in
hand-written HIP kernels for gfx950 behind the reference's Python API.
"""

__version__ = "1.0.0"
