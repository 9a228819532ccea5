"""getdp_amd -- MI355X (gfx950) backend for GetDP's finite-element hot path:
batched Galerkin assembly and the Sparskit-equivalent Krylov solve, as
hand-written HIP kernels behind a C-ABI (include/gdp_hip.h).

Python here is a thin mirror of the reference's LinAlg interface for tests
and benchmarks; the work runs in libgdp_hip.so.
"""
from . import problems  # noqa: F401
from ._lib import GdpError, LIB_PATH, lib  # noqa: F401

__all__ = ["problems", "lib", "LIB_PATH", "GdpError"]
