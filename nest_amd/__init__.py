"""nest_amd — MI355X (gfx950) HIP backend for nest's Euler finite-volume step.

The compute lives in ``libnest_hip.so`` (hand-written HIP kernels behind the C
ABI of ``include/nest_hip.h``); this package is the Python front end.
"""
from ._lib import NestError, NestLibraryMissing, load  # noqa: F401
from .euler import Sim2D, Solver1D  # noqa: F401

__all__ = ["Sim2D", "Solver1D", "NestError", "NestLibraryMissing", "load"]
