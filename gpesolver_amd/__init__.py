"""gpesolver_amd -- MI355X (gfx950) backend for GPESolver's hot path.

The compute path is libgpehip.so (hand-written HIP kernels behind the C ABI
of include/gpehip.h); this package holds its ctypes binding (hip), the host
configuration mirror (params) and the solver front end (solver).
"""
from . import hip, params  # noqa: F401

__all__ = ["hip", "params"]
