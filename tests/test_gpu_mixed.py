"""Mixed precision: FP32 V-cycles as the preconditioner of the FP64 CG (the
FP64 defect-correction loop keeps the residual, the reductions, the update and
the coarsest solve in FP64), and the FP32-vector CG whose FP32 recursion is
restarted from the FP64 residual by reliable updates.  Each solve must reach
the same true FP64 relative residual as the plain FP64 method, in about as
many cycles, and the per-field error against the FP64 oracle is reported."""
import numpy as np
import pytest

from helpers import pair, params, rel_err
from impec2d_amd import fields

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("mode", ["mixed", "mixed32"])
@pytest.mark.parametrize("n1,n2,aniso", [(256, 256, False), (200, 120, False), (77, 45, False),
                                         (128, 128, True), (1024, 1024, False)])
def test_mixed_pressure_matches_fp64(gpu_lib, n1, n2, aniso, mode):
    from impec2d_amd import capi
    prec_m = capi.PREC_MIXED if mode == "mixed" else capi.PREC_MIXED32
    p = params(N1=n1, N2=n2, nu=1.0, eps=1e-10, mu_inj=1.0 / 20)
    K11 = fields.lognormal_permeability(n1, n2, sigma=1.0, corr_len=0.05, seed=1)
    K22 = fields.lognormal_permeability(n1, n2, sigma=1.0, corr_len=0.05, seed=2) if aniso else None
    _say("mixed pressure", mode, n1, n2, aniso)
    out = {}
    for prec in (capi.PREC_FP64, prec_m):
        o, g = pair(p, K11, K22, precision=prec)
        rng = np.random.default_rng(5)
        c = np.clip(rng.random(o.scalar().shape), 0, 1)
        g.set_concentration(c)
        g.set_source(o.source("one_cell"))
        its, rr = g.pressure_solve()
        assert rr <= 1e-10, (prec, rr)
        out[prec] = (its, o.interior(g.get_pressure()))
        g.close()
    its64, p64 = out[capi.PREC_FP64]
    its32, p32 = out[prec_m]
    _say("  its fp64 %d %s %d, |p32-p64|/|p64| %.2e" % (its64, mode, its32, rel_err(p32, p64)))
    # anisotropic K: the coarse levels coarsen the x-mobility only, the FP64
    # cycle needs ~100 iterations and FP32 rounding adds a few
    assert its32 <= (its64 * 1.1 + 2 if aniso else its64 + 2)
    assert rel_err(p32, p64) < 1e-8


@pytest.mark.parametrize("mode", ["mixed", "mixed32"])
@pytest.mark.parametrize("kw", [dict(), dict(diffusion=1, alpha_abs=1, alpha_l=1e-3, alpha_t=1e-4, d_m=1e-5)])
def test_mixed_steps_error_per_field(gpu_lib, kw, mode):
    """Ten IMPEC steps with FP32 V-cycles against the FP64 oracle, pressure
    solved to 1e-12: every field within 1e-10 relative, the same bar as the FP64
    method (the preconditioner's precision does not limit the FP64 residual)."""
    from impec2d_amd import capi
    n = 64
    p = params(**dict(dict(N1=n, N2=n, eps=1e-12, para_eps=1e-13, mu_inj=1.0 / 20), **kw))
    K = fields.lognormal_permeability(n, n, sigma=1.0, corr_len=0.05, seed=7)
    o, g = pair(p, K, None, precision=capi.PREC_MIXED if mode == "mixed" else capi.PREC_MIXED32)
    q = o.source("one_cell")
    c, pr, v = o.scalar(), o.scalar(), o.velocity()
    g.set_source(q)
    g.set_concentration(c)
    t_ref = t_gpu = 0.0
    for _ in range(10):
        st, dt = o.step(c, pr, v, q, t_ref)
        assert st == 0
        t_ref += dt
        t_gpu += g.step(t_gpu)
    ep = rel_err(o.interior(g.get_pressure()), o.interior(pr))
    ev = rel_err(g.get_velocity(), v)
    ec = rel_err(o.interior(g.get_concentration()), o.interior(c))
    _say("%s vs FP64 oracle after 10 steps: p %.2e v %.2e c %.2e" % (mode, ep, ev, ec))
    assert ep < 1e-10 and ev < 1e-10 and ec < 1e-10


TRANSPORT32_TOL = 1e-5


@pytest.mark.parametrize("kw", [dict(), dict(diffusion=1, alpha_abs=1, alpha_l=1e-3, alpha_t=1e-4, d_m=1e-5)])
def test_fp32_transport_error_per_field(gpu_lib, kw):
    """The mixed-precision configuration in full: FP32 V-cycles in the FP64 CG
    and the parallel-splitting transport with FP32 reconstructions, half steps
    and fluxes (states, updates and mass sums FP64), ten steps on a lognormal-K
    quarter-five-spot (M = 20) against the FP64 oracle.  The error per field is
    stated (within TRANSPORT32_TOL relative: FP32 arithmetic, about 1e-7 per
    step); the FP32 path is really taken (c differs from the FP64 run); the
    mass balance stays at FP64 level."""
    from impec2d_amd import capi
    n = 64
    p = params(**dict(dict(N1=n, N2=n, eps=1e-12, para_eps=1e-13, mu_inj=1.0 / 20), **kw))
    K = fields.lognormal_permeability(n, n, sigma=1.0, corr_len=0.05, seed=7)
    o, g = pair(p, K, None, precision=capi.PREC_MIXED, transport32=1)
    _, g64 = pair(p, K, None, precision=capi.PREC_MIXED)
    q = o.source("one_cell")
    c, pr, v = o.scalar(), o.scalar(), o.velocity()
    for h in (g, g64):
        h.set_source(q)
        h.set_concentration(c)
    t_ref = t_gpu = t64 = 0.0
    for _ in range(10):
        st, dt = o.step(c, pr, v, q, t_ref)
        assert st == 0
        t_ref += dt
        t_gpu += g.step(t_gpu)
        t64 += g64.step(t64)
    ep = rel_err(o.interior(g.get_pressure()), o.interior(pr))
    ev = rel_err(g.get_velocity(), v)
    ec = rel_err(o.interior(g.get_concentration()), o.interior(c))
    e64 = rel_err(o.interior(g64.get_concentration()), o.interior(c))
    me = g.mass_error(t_gpu)
    _say("FP32 transport vs FP64 oracle after 10 steps: p %.2e v %.2e c %.2e (FP64 transport: c %.2e); "
         "mass error %.2e" % (ep, ev, ec, e64, me))
    assert ep < TRANSPORT32_TOL and ev < TRANSPORT32_TOL and ec < TRANSPORT32_TOL
    assert ec > 10 * e64            # the FP32 sweeps ran
    assert me < 1e-9
