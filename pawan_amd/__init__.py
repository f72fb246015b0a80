"""MI355X (gfx950) implementation of PAWAN's particle-interaction and
time-integration hot path.

The product is libpawan_gpu.so (hand-written HIP kernels behind the C-ABI of
include/pawan_gpu.h, plus the C++ class API in csrc/host/pawan/).  This
package holds its sources, the build script, a ctypes binding (capi) and the
synthetic workload generators (workloads).
"""
from . import workloads  # noqa: F401

__all__ = ["workloads", "capi", "build"]
