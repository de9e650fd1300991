"""Failure handling of the C ABI, forced from the host: every kernel still
exits normally, and the entry points report the failure through their status
and impec_last_error() instead of returning a result that misses its
contract.  The reference ignores non-convergence (MGSolver.cpp:247 returns
true whatever HYPRE reports); this library returns IMPEC_ENOCONV / IMPEC_ENAN."""
import numpy as np
import pytest

from tests.helpers import pair, params, random_c, smooth_c

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _code(excinfo):
    return excinfo.value.code


def test_pressure_iteration_limit_returns_enoconv(gpu_lib):
    """max_iterations = 3 with a tolerance no solve reaches in 3 iterations."""
    from impec2d_amd import capi
    _say("pressure: iteration limit")
    p = params(N1=64, N2=64, eps=1e-14, mu_inj=0.1)
    o, g = pair(p, max_iterations=3)
    g.set_concentration(smooth_c(o, seed=3))
    g.set_source(o.source("one_cell"))
    with pytest.raises(capi.ImpecError) as e:
        g.pressure_solve()
    assert _code(e) == capi.IMPEC_ENOCONV
    assert "no convergence after 3 iterations" in str(e.value)
    assert g.stats().last_its == 3


@pytest.mark.parametrize("nu", [1.0, 2 / 3.0])
def test_pressure_tolerance_below_rounding_floor_converges_at_floor(gpu_lib, nu):
    """eps = 1e-18 lies below what ||b - A p|| / ||b|| can resolve in FP64: the
    recursion's residual falls below it, the true residual cannot.  The solve
    stops at the rounding floor of its own residual evaluation (reported as
    last_relres_floor) instead of iterating to its limit, and p equals a solve
    at an attainable tolerance."""
    from impec2d_amd import capi
    _say("pressure: eps below the rounding floor, nu", nu)
    res = {}
    for eps in (1e-18, 1e-12):
        p = params(N1=48, N2=48, eps=eps, mu_inj=0.1, nu=nu)
        o, g = pair(p, max_iterations=2000)
        g.set_concentration(smooth_c(o, seed=4))
        g.set_source(o.source("one_cell"))
        its, rr = g.pressure_solve()
        st = g.stats()
        _say("  eps %.0e: %d iterations, relres %.2e, floor %.2e" % (eps, its, rr, st.last_relres_floor))
        assert rr <= max(eps, st.last_relres_floor)
        assert 1e-16 < st.last_relres_floor < 1e-11
        res[eps] = (its, o.interior(g.get_pressure()))
    assert res[1e-18][0] < 200          # no run to the iteration limit
    from tests.helpers import rel_err
    assert rel_err(res[1e-18][1], res[1e-12][1]) < 1e-10


def test_stalled_recursion_stops_at_floor(gpu_lib):
    """The production configuration parallel.cfg (141^2, circle wells, M = 41,
    nu = 2/3) with its eps = 1e-14, below the rounding floor of its systems
    (about 3e-12): now and then the recursion's residual stops falling short of
    1e-14.  Without a stall test such a solve runs to the iteration limit while
    its iterate drifts (true residual 1e-8 after 10000 iterations); with it,
    every one of 40 steps ends at max(eps, floor) in a bounded number of
    iterations."""
    import json
    import os
    from oracle import pyoracle as po
    from tests.helpers import gpu_params
    from impec2d_amd import capi
    _say("pressure: stalled recursion at the floor (parallel.cfg workload, 40 steps)")
    fix = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "longrun", "parallel_cfg.npz")
    meta = json.loads(str(np.load(fix)["meta"]))
    P = po.default_params(**dict(meta["params"], eps=1e-14))
    o = po.Oracle(P)
    g = capi.Context(gpu_params(P))
    g.set_source(o.source(meta["profile"]))
    g.set_concentration(o.scalar())
    t, worst, restarts = 0.0, 0, 0
    for _ in range(40):
        t += g.step(t, 0.4)
        st = g.stats()
        assert st.last_relres <= max(1e-14, st.last_relres_floor)
        worst = max(worst, st.last_its)
        restarts += st.last_restarts
    _say("  most iterations in one solve %d, restarts %d, last relres %.2e floor %.2e"
         % (worst, restarts, st.last_relres, st.last_relres_floor))
    assert worst <= 120
    g.close()


def test_nan_concentration_returns_enan(gpu_lib):
    """A NaN in c poisons the mobility, hence the operator and the residual."""
    from impec2d_amd import capi
    _say("pressure: NaN in c")
    p = params(N1=32, N2=32, eps=1e-10, mu_inj=0.1)
    o, g = pair(p)
    c = smooth_c(o, seed=5)
    o.interior(c)[10, 17] = np.nan
    g.set_concentration(c)
    g.set_source(o.source("one_cell"))
    with pytest.raises(capi.ImpecError) as e:
        g.step(0.0)
    assert _code(e) == capi.IMPEC_ENAN
    assert "non-finite" in str(e.value)


def _dispersion_case(**kw):
    p = params(**dict(dict(N1=40, N2=40, eps=1e-13, para_eps=1e-10, mu_inj=0.1, diffusion=1, alpha_abs=1,
                           alpha_l=1e-2, alpha_t=1e-3, d_m=1e-4), **kw))
    return p


def test_dispersion_iteration_limit_returns_enoconv(gpu_lib):
    from impec2d_amd import capi
    _say("dispersion: iteration limit")
    p = _dispersion_case(para_eps=1e-15)
    o, g = pair(p, max_iterations=2)
    c = random_c(o, seed=6)
    v = o.velocity()
    rng = np.random.default_rng(7)
    v[:] = rng.uniform(-50, 50, v.shape)
    g.set_concentration(c)
    g.set_velocity(v)
    with pytest.raises(capi.ImpecError) as e:
        g.dispersion_cn(1e-3)
    assert _code(e) == capi.IMPEC_ENOCONV
    assert "dispersion solve: no convergence" in str(e.value)


def test_dispersion_reports_true_residual(gpu_lib):
    """The BiCGStab recursion's residual is rebuilt from partial sums and can
    cancel; the solve confirms convergence on ||b - A x|| itself and reports
    that value, which must agree with the oracle's matrix applied to the
    returned solution."""
    _say("dispersion: true residual")
    p = _dispersion_case()
    o, g = pair(p)
    c = random_c(o, seed=8)
    v = o.velocity()
    rng = np.random.default_rng(9)
    v[:] = rng.uniform(-50, 50, v.shape)
    o.fill_velocity(v)
    g.set_concentration(c)
    g.set_velocity(v)
    dt = 2e-3
    A, b = g.parabolic_operator(dt)
    g.set_concentration(c)
    its = g.dispersion_cn(dt)
    x = o.interior(g.get_concentration()).ravel()
    from oracle import pyoracle as po
    r = b - po.matvec9(o.N1, o.N2, A, x)
    rel = np.linalg.norm(r) / np.linalg.norm(b)
    st = g.stats()
    _say("  its %d, reported %.3e, recomputed %.3e" % (its, st.last_parabolic_relres, rel))
    assert rel <= 1e-10
    assert st.last_parabolic_relres <= 1e-10
    assert abs(st.last_parabolic_relres - rel) <= 1e-3 * rel + 1e-15


def test_dispersion_tolerance_below_rounding_floor_returns_enoconv(gpu_lib):
    from impec2d_amd import capi
    _say("dispersion: para_eps below the rounding floor")
    p = _dispersion_case(para_eps=1e-19)
    o, g = pair(p, max_iterations=500)
    c = random_c(o, seed=10)
    v = o.velocity()
    rng = np.random.default_rng(11)
    v[:] = rng.uniform(-50, 50, v.shape)
    g.set_concentration(c)
    g.set_velocity(v)
    with pytest.raises(capi.ImpecError) as e:
        g.dispersion_cn(2e-3)
    assert _code(e) == capi.IMPEC_ENOCONV
    assert g.stats().last_parabolic_relres < 1e-12


@pytest.mark.parametrize("layers", [4, 8])
def test_garbage_network_still_reaches_tolerance(gpu_lib, layers):
    """A learned guess from a network whose weights are scaled by 1e6 (with 8
    layers the BF16 activations overflow and the increment is not finite): the
    minimal-residual step lengths discard the increment and the MG-CG solve
    still reaches the tolerance (plain multigrid is the fallback and decides convergence)."""
    from impec2d_amd import capi
    _say("learned guess: garbage network, layers", layers)
    n = 64
    p = params(N1=n, N2=n, eps=1e-10, mu_inj=0.1)
    net = capi.Network(layers=layers, capacity=16, crop=32, seed=2)
    w = net.params() * np.float32(1e6)
    net.set_params(w)
    o, g = pair(p, initial_guess=capi.GUESS_LEARNED)
    g.set_warmstart(net)
    g.set_source(o.source("one_cell"))
    g.set_concentration(o.scalar())
    t = 0.0
    for _ in range(6):
        t += g.step(t)
        st = g.stats()
        assert st.last_relres <= 1e-10
        assert np.isfinite(st.last_guess_alpha)
    g.close()
    net.close()
