"""impec2d_amd: the IMPEC hot path of Impec2D on AMD MI355X (gfx950).

The compute lives in libimpec_hip.so (hand-written HIP kernels behind the C
ABI of include/impec_hip.h); :mod:`impec2d_amd.capi` binds it with ctypes.  The C++ solver classes that
restate Impec2DSolver and its solvers over the same ABI live in
``impec2d_amd/host`` (libimpec_host.so and the impec2d_run driver).
"""
from . import capi  # noqa: F401

__all__ = ["capi"]
