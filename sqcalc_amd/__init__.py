"""MI355X (gfx950) reciprocal-space S(q) path: a drop-in for the GPU plug-in of sqcalc.

The per-frame work runs in libsqc_hip.so (hand-written HIP kernels, C ABI in
include/sqc_hip.h); this package holds the ctypes binding, the host-side
mirror of the reference plug-in interface and the autograd entry point.
"""

from . import _lib  # noqa: F401
from .structure_factor import HipStructureFactor, cell_from_bounds  # noqa: F401

__all__ = ["HipStructureFactor", "cell_from_bounds"]
