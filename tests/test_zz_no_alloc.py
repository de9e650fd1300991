"""impec_max_stable_timestep_noalloc (the entry point that allocates nothing)
against impec_max_stable_timestep on the same velocities: the same time step
bit for bit (max / min reductions are exact), on two calls in a row, the
second on a smaller grid, in the CFL modes of the reference (combined
advective, advection-only test mode, explicit-dispersion limit).  Device
buffers for the caller's side come from the HIP runtime the library itself
uses; the entry point receives them and allocates nothing."""
import ctypes as C

import numpy as np
import pytest

from tests.helpers import gpu_params, params

pytestmark = pytest.mark.gpu


def _hip():
    hip = C.CDLL("libamdhip64.so.7")   # already loaded with libimpec_hip.so
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipDeviceSynchronize.argtypes = []
    return hip


H2D, D2H = 1, 2


def _dev_copy(hip, a):
    p = C.c_void_p()
    assert hip.hipMalloc(C.byref(p), a.nbytes) == 0
    assert hip.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, H2D) == 0
    return p


CASES = [
    dict(),                                                              # advective CFL (CN dispersion or none)
    dict(is_test=1),                                                     # MaxStableAdvectionTimestep
    dict(diffusion=1, para_cn=0, alpha_l=1e-2, alpha_t=2e-3),            # explicit-dispersion limit
]


@pytest.mark.parametrize("kw", CASES)
def test_noalloc_timestep_equals_context_timestep(gpu_lib, kw):
    from impec2d_amd import capi
    hip = _hip()
    rng = np.random.default_rng(5)
    for n1, n2 in ((200, 136), (37, 21)):                # the second call on a smaller grid
        print("noalloc", kw, n1, n2, flush=True)
        p = params(**dict(dict(N1=n1, N2=n2, x_ub=1.3, y_ub=0.9, cfl=0.45), **kw))
        g = gpu_params(p)
        ctx = capi.Context(g)
        v = rng.uniform(-2.0, 2.0, ctx.velocity_shape())
        ctx.set_velocity(v)
        dt_ref = ctx.max_stable_timestep()
        ctx.close()
        vd = _dev_copy(hip, np.ascontiguousarray(v))
        dtd = _dev_copy(hip, np.array([-1.0]))
        st = gpu_lib.impec_max_stable_timestep_noalloc(C.byref(g), vd, dtd, None)
        assert st == 0, gpu_lib.impec_last_error(None)
        out = np.zeros(1)
        assert hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), dtd, 8, D2H) == 0
        hip.hipFree(vd)
        hip.hipFree(dtd)
        assert out[0] == dt_ref, (out[0], dt_ref)


def test_noalloc_rejects_bad_sizes(gpu_lib):
    from impec2d_amd import capi
    g = gpu_params(params(N1=70000, N2=70000))
    st = gpu_lib.impec_max_stable_timestep_noalloc(C.byref(g), C.c_void_p(8), C.c_void_p(8), None)
    assert st == capi.IMPEC_EINVAL
    assert b"exceeds the maximum" in gpu_lib.impec_last_error(None)
    g = gpu_params(params(N1=0, N2=16))
    assert gpu_lib.impec_max_stable_timestep_noalloc(C.byref(g), C.c_void_p(8), C.c_void_p(8), None) == capi.IMPEC_EINVAL
