"""MI355X-native wavefront path of AtmosphericTurbulenceSimul.

libatsim.so (built by `python -m atmosphericturbulencesimul_amd.build`)
holds the hand-written gfx950 HIP kernels and the C ABI (include/atsim.h);
this package is its Python front end.  There is no CPU fallback.
"""
from ._lib import AtsimError, LIB_PATH, lib  # noqa: F401

__all__ = ["AtsimError", "LIB_PATH", "lib"]
