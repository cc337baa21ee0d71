"""pulse_amd: the path-tracing hot path of pulse (TLAS/BLAS traversal, GGX,
next event estimation, accumulation) as HIP kernels for MI355X (gfx950)
behind a C ABI (include/pulse_hip.h).  The Python modules are a thin host
layer over libpulse_hip.so."""
from . import _native  # noqa: F401
