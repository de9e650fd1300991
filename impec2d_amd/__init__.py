"""impec2d_amd: the IMPEC hot path of Impec2D on AMD MI355X (gfx950).

The compute lives in libimpec_hip.so (hand-written HIP kernels behind the C
ABI of include/impec_hip.h); :mod:`impec2d_amd.capi` binds it with ctypes and
:mod:`impec2d_amd.solver` restates the Impec2DSolver time loop on top of it.
"""
from . import capi  # noqa: F401

__all__ = ["capi"]
