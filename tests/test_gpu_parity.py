"""Parity of the HIP path against the CPU oracle on the same inputs.

Per-cell arithmetic (operator assembly, velocities, CFL, sources, transport
sweeps, dispersion operator) is compared bit for bit; linear solves are
compared to the solver tolerance (relative to max|field|).
"""
import numpy as np
import pytest

from tests.helpers import pair, params, random_c, rel_err, smooth_c

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


OP_CASES = [
    dict(nu=1.0),
    dict(nu=2 / 3.0),
    dict(nu=0.0),
    dict(nu=2 / 3.0, avgproc=1),
    dict(nu=2 / 3.0, stype=0, x_ub=2 ** 0.5, y_ub=2 ** 0.5, mu_res=41.0, mu_inj=1.0),
    dict(nu=2 / 3.0, N1=24, N2=12),          # dx != dy
    dict(nu=1.0, N1=17, N2=9),               # odd sizes
]


@pytest.mark.parametrize("kw", OP_CASES)
def test_pressure_operator_bitexact(gpu_lib, kw):
    p = params(**dict(dict(N1=16, N2=16), **kw))
    _say("operator", kw)
    o, g = pair(p)
    c = random_c(o, seed=1)
    g.set_concentration(c)
    A_gpu = g.pressure_operator()
    A_ref = o.assemble(c)
    assert np.array_equal(A_gpu, A_ref)


def test_pressure_operator_bitexact_anisotropic_K(gpu_lib):
    p = params(N1=12, N2=12, nu=2 / 3.0)
    o0 = __import__("oracle.pyoracle", fromlist=["x"]).Oracle(p)
    rng = np.random.default_rng(5)
    K11 = np.exp(rng.standard_normal(o0.scalar().shape))
    K22 = np.exp(rng.standard_normal(o0.scalar().shape))
    _say("operator anisotropic K")
    o, g = pair(p, K11, K22)
    c = random_c(o, seed=2)
    g.set_concentration(c)
    assert np.array_equal(g.pressure_operator(), o.assemble(c))


@pytest.mark.parametrize("kw", [dict(nu=1.0), dict(nu=2 / 3.0), dict(nu=0.3, avgproc=1),
                                dict(nu=2 / 3.0, N1=20, N2=13)])
def test_velocities_bitexact(gpu_lib, kw):
    p = params(**dict(dict(N1=16, N2=16), **kw))
    _say("velocities", kw)
    o, g = pair(p)
    c = random_c(o, seed=3)
    rng = np.random.default_rng(4)
    pr = o.scalar()
    o.interior(pr)[:] = rng.standard_normal((o.N1, o.N2))
    g.set_concentration(c)
    g.set_pressure(pr)
    g.velocities()
    v_ref = o.velocity()
    pr_ref = pr.copy()
    o.velocities(pr_ref, v_ref, c)
    v_gpu = g.get_velocity()
    b = o.b
    # every entry the reference writes (the two never-written rows stay 0 on both sides)
    assert np.array_equal(v_gpu, v_ref)
    assert np.array_equal(g.get_pressure(), pr_ref)
    # and the time step derived from them
    assert g.max_stable_timestep() == o.timestep(v_ref)
    del b


def test_timestep_modes_bitexact(gpu_lib):
    rng = np.random.default_rng(7)
    for kw in [dict(), dict(is_test=1), dict(diffusion=1, para_cn=0, alpha_l=0.3, alpha_t=0.1)]:
        p = params(N1=18, N2=14, **kw)
        _say("timestep", kw)
        o, g = pair(p)
        v = rng.standard_normal(o.velocity().shape)
        g.set_velocity(v)
        assert g.max_stable_timestep() == o.timestep(v)


@pytest.mark.parametrize("profile,stype", [("one_cell", 1), ("one_cell", 0), ("circle", 1), ("circle", 0)])
def test_source_split_bitexact(gpu_lib, profile, stype):
    p = params(N1=32, N2=32, stype=stype, well_rc=0.2)
    _say("source", profile, stype)
    o, g = pair(p)
    q = o.source(profile)
    c = random_c(o, seed=8)
    g.set_source(q)
    g.set_concentration(c)
    dt = 1.7e-4
    for _ in range(3):
        g.source_split(dt)
        o.L.or_source_split(o.ctx, c, q, dt)
    cg = g.get_concentration()
    assert np.array_equal(o.interior(cg), o.interior(c))
    out_gpu, out_ref = g.stats().total_mass_outflux, o.total_mass_outflux
    if profile == "one_cell":
        assert out_gpu == out_ref          # one or two sinks: same summation order
    else:
        assert abs(out_gpu - out_ref) <= 1e-14 * abs(out_ref)


@pytest.mark.parametrize("kw", [dict(dim_split=4), dict(dim_split=0), dict(dim_split=1), dict(dim_split=2),
                                dict(dim_split=3), dict(dim_split=4, hyper=1), dict(dim_split=4, porosity=0.7),
                                dict(dim_split=4, is_test=1), dict(dim_split=4, N1=21, N2=11),
                                # the one-pass parallel splitting: widths across its 60-column
                                # strips, ragged row chunks, tiny grids
                                dict(dim_split=4, N1=7, N2=61), dict(dim_split=4, N1=70, N2=121, hyper=1),
                                dict(dim_split=4, N1=3, N2=2), dict(dim_split=4, N1=130, N2=57, porosity=0.7)])
def test_advect_split_bitexact(gpu_lib, kw):
    p = params(**dict(dict(N1=24, N2=24), **kw))
    _say("advect", kw)
    o, g = pair(p)
    c = smooth_c(o, seed=9)
    rng = np.random.default_rng(10)
    v = rng.standard_normal(o.velocity().shape)
    g.set_concentration(c)
    g.set_velocity(v)
    dt = 0.2 / max(o.N1, o.N2) / np.abs(v).max()
    for _ in range(2):
        g.advect_split(dt)
        o.advect_split(c, v, dt)
    cg = g.get_concentration()
    assert np.array_equal(o.interior(cg), o.interior(c))


@pytest.mark.parametrize("kw", [dict(), dict(hyper=1), dict(N1=13, N2=21, porosity=0.7),
                                dict(N1=9, N2=5, hyper=1, porosity=0.5)])
def test_intercell_flux_bitexact(gpu_lib, kw):
    """CalculateIntercellFlux_F/G (MUSCL and upwind) on the device equal the
    oracle's fluxes bit for bit, ghost rows and columns included."""
    p = params(**dict(dict(N1=24, N2=24), **kw))
    _say("intercell flux", kw)
    o, g = pair(p)
    rng = np.random.default_rng(21)
    c = rng.random(o.scalar().shape)
    v = rng.standard_normal(o.velocity().shape)
    g.set_concentration(c)
    g.set_velocity(v)
    dt = 0.3 / max(o.N1, o.N2) / np.abs(v).max()
    for d in (0, 1):
        assert np.array_equal(g.intercell_flux(d, dt), o.flux(d, c, v, dt)), d


@pytest.mark.parametrize("kw", [dict(), dict(alpha_abs=1, alpha_l=1e-3, alpha_t=1e-4, d_m=1e-5),
                                dict(N1=18, N2=10, porosity=0.8)])
def test_parabolic_operator_bitexact(gpu_lib, kw):
    p = params(**dict(dict(N1=16, N2=16, diffusion=1, alpha_l=0.5, alpha_t=0.2), **kw))
    _say("parabolic operator", kw)
    o, g = pair(p)
    c = smooth_c(o, seed=11)
    rng = np.random.default_rng(12)
    v = rng.standard_normal(o.velocity().shape) + 0.5
    o.fill_velocity(v)
    g.set_concentration(c)
    g.set_velocity(v)
    dt = 3e-3
    A, r = g.parabolic_operator(dt)
    assert np.array_equal(A, o.parabolic_matrix(v, dt))
    assert np.array_equal(r, o.parabolic_rhs(c, v, dt))


def test_parabolic_solve_matches_dense(gpu_lib):
    from oracle import pyoracle as po
    p = params(N1=16, N2=16, diffusion=1, alpha_l=0.5, alpha_t=0.2, para_eps=1e-14)
    _say("parabolic solve")
    o, g = pair(p)
    c = smooth_c(o, seed=13)
    rng = np.random.default_rng(14)
    v = rng.standard_normal(o.velocity().shape) + 0.5
    o.fill_velocity(v)
    g.set_concentration(c)
    g.set_velocity(v)
    dt = 3e-3
    its = g.dispersion_cn(dt)
    x = po.dense_solve9(o.N1, o.N2, o.parabolic_matrix(v, dt), o.parabolic_rhs(c, v, dt))
    cg = o.interior(g.get_concentration())
    assert its > 0
    assert rel_err(cg.reshape(-1), x) < 1e-12


@pytest.mark.parametrize("n1,n2", [(5, 3), (9, 61), (21, 63), (13, 125), (33, 62)])
@pytest.mark.parametrize("cn_guess", [0, 1])
def test_parabolic_solve_ragged_strips(gpu_lib, n1, n2, cn_guess):
    """The streamed BiCGStab on grids whose widths cross the 60- and
    62-column strips of its launches and whose heights leave ragged row
    chunks (one-row chunks included), from the zero guess (the one-pass
    start) and from the current concentration, matches the dense solve of
    the oracle's operator to 1e-12."""
    from oracle import pyoracle as po
    p = params(N1=n1, N2=n2, diffusion=1, alpha_l=0.5, alpha_t=0.2, para_eps=1e-14)
    _say("parabolic solve ragged", n1, n2, cn_guess)
    o, g = pair(p, cn_guess=cn_guess)
    c = smooth_c(o, seed=21)
    rng = np.random.default_rng(22)
    v = rng.standard_normal(o.velocity().shape) + 0.5
    o.fill_velocity(v)
    g.set_concentration(c)
    g.set_velocity(v)
    dt = 3e-3
    its = g.dispersion_cn(dt)
    x = po.dense_solve9(o.N1, o.N2, o.parabolic_matrix(v, dt), o.parabolic_rhs(c, v, dt))
    cg = o.interior(g.get_concentration())
    assert its > 0
    assert rel_err(cg.reshape(-1), x) < 1e-12


@pytest.mark.parametrize("kw", [dict(nu=1.0), dict(nu=2 / 3.0), dict(nu=2 / 3.0, stype=0, x_ub=2 ** 0.5,
                                                                      y_ub=2 ** 0.5),
                                dict(nu=2 / 3.0, N1=24, N2=12), dict(nu=1.0, avgproc=1)])
def test_pressure_solve_matches_dense(gpu_lib, kw):
    p = params(**dict(dict(N1=24, N2=24, eps=1e-13, mu_inj=0.1), **kw))
    _say("pressure solve", kw)
    o, g = pair(p)
    c = smooth_c(o, seed=15)
    q = o.source("one_cell")
    g.set_concentration(c)
    g.set_source(q)
    its, relres = g.pressure_solve()
    p_ref = o.scalar()
    assert o.pressure_solve_dense(c, p_ref, q) == 0
    pg = g.get_pressure()
    assert relres <= 1e-13
    assert rel_err(o.interior(pg), o.interior(p_ref)) < 1e-11
    assert its < 60


def test_pressure_solve_circle_wells(gpu_lib):
    p = params(N1=32, N2=32, eps=1e-13, mu_inj=0.1, well_rc=0.1, nu=2 / 3.0)
    _say("pressure solve circle")
    o, g = pair(p)
    c = smooth_c(o, seed=16)
    q = o.source("circle")
    g.set_concentration(c)
    g.set_source(q)
    its, relres = g.pressure_solve()
    p_ref = o.scalar()
    assert o.pressure_solve_dense(c, p_ref, q) == 0
    assert relres <= 1e-13
    assert rel_err(o.interior(g.get_pressure()), o.interior(p_ref)) < 1e-11


STEP_CASES = [
    dict(nu=1.0),
    dict(nu=2 / 3.0),
    dict(nu=1.0, stype=0, x_ub=2 ** 0.5, y_ub=2 ** 0.5),
    dict(nu=2 / 3.0, diffusion=1, alpha_abs=1, alpha_l=1e-3, alpha_t=1e-4, d_m=1e-5),
    dict(nu=2 / 3.0, diffusion=1, alpha_l=1.0, alpha_t=1.0, mu_res=41.0, mu_inj=1.0, stype=0,
         x_ub=2 ** 0.5, y_ub=2 ** 0.5),
]


@pytest.mark.parametrize("kw", STEP_CASES)
def test_steps_1_and_10(gpu_lib, kw):
    """Pressure, velocity and concentration after 1 and 10 IMPEC steps agree
    with the oracle to 1e-10 relative (FP64, both solves at 1e-13)."""
    p = params(**dict(dict(N1=32, N2=32, eps=1e-13, para_eps=1e-13, mu_inj=0.1), **kw))
    _say("steps", kw)
    o, g = pair(p)
    q = o.source("one_cell")
    c, pr, v = o.scalar(), o.scalar(), o.velocity()
    g.set_source(q)
    g.set_concentration(c)
    t_ref = t_gpu = 0.0
    for k in range(1, 11):
        st, dt_ref = o.step(c, pr, v, q, t_ref)
        assert st == 0
        t_ref += dt_ref
        dt_gpu = g.step(t_gpu)
        t_gpu += dt_gpu
        if k in (1, 10):
            cg, pg, vg = g.get_concentration(), g.get_pressure(), g.get_velocity()
            assert rel_err(o.interior(pg), o.interior(pr)) < 1e-10, k
            assert rel_err(vg, v) < 1e-10, k
            assert rel_err(o.interior(cg), o.interior(c)) < 1e-10, k
            assert abs(dt_gpu - dt_ref) <= 1e-10 * dt_ref
            # diagnostics of the same states (Impec2DSolver.cpp:530-620): mass
            # and produced mass to 1e-10 relative, the balance and the
            # divergence / wall residuals (rounding-level numbers) absolutely
            assert abs(g.total_mass() - o.mass(c)) <= 1e-10 * o.mass(c)
            # produced mass: relative to itself once it is more than a trace of
            # the injected mass (dispersion carries traces of 1e-90 to the sink)
            tmo_g, tmo_o = g.stats().total_mass_outflux, o.total_mass_outflux
            injected = t_ref * p["I"] / (o.dx * o.dy)
            assert abs(tmo_g - tmo_o) <= 1e-10 * max(tmo_o, 1e-12 * injected)
            assert abs(g.mass_error(t_gpu) - o.mass_error(c, t_ref)) <= 1e-11
            dq = 2.0 * p["I"]   # sum |q| dx dy of the two wells
            assert abs(g.divergence_error() - o.div_error(v, q)) <= 1e-9 * dq
            assert abs(g.velocity_boundary_error() - o.noflow_error(v)) <= 1e-12


@pytest.mark.parametrize("kw", [dict(nu=1.0), dict(nu=2 / 3.0, stype=0, x_ub=2 ** 0.5, y_ub=2 ** 0.5),
                                dict(nu=1.0, N1=37, N2=21)])
def test_diagnostics_match_oracle(gpu_lib, kw):
    """impec_total_mass, impec_mass_error, impec_divergence_error and
    impec_velocity_boundary_error (Impec2DSolver.cpp:530-620) against the
    oracle's values on a state where none of them is near zero: random c,
    random face velocities including the wall faces, the well sources."""
    p = params(**dict(dict(N1=32, N2=32), **kw))
    _say("diagnostics", kw)
    o, g = pair(p)
    c = random_c(o, seed=21)
    rng = np.random.default_rng(22)
    v = rng.uniform(-3.0, 3.0, o.velocity().shape)
    q = o.source("one_cell")
    g.set_concentration(c)
    g.set_velocity(v)
    g.set_source(q)
    m_ref = o.mass(c)
    assert m_ref > 0
    assert abs(g.total_mass() - m_ref) <= 1e-13 * m_ref
    for t in (0.37, 1.25):
        e_ref = o.mass_error(c, t)
        assert e_ref > 1e-3
        assert abs(g.mass_error(t) - e_ref) <= 1e-13 * e_ref
    d_ref = o.div_error(v, q)
    assert d_ref > 1e-3
    assert abs(g.divergence_error() - d_ref) <= 1e-13 * d_ref
    w_ref = o.noflow_error(v)
    assert w_ref > 1e-3
    assert abs(g.velocity_boundary_error() - w_ref) <= 1e-13 * w_ref


@pytest.mark.parametrize("n1,n2", [(256, 256), (200, 120), (77, 45), (2048, 2048)])
def test_fused_and_unfused_multigrid_agree(gpu_lib, n1, n2):
    """The streaming V-cycle (with its single-workgroup bottom kernel), the same
    legs run down to the coarsest level, and the one-kernel-per-pass V-cycle are
    preconditioners of the same CG; all reach the tolerance and agree on p, and
    the bottom kernel does what the small-level legs do (same iteration count)."""
    from impec2d_amd import capi
    # 1e-11: the FP64 rounding floor of ||b - A p|| / ||b|| at 2048^2 is ~1e-12
    p = params(N1=n1, N2=n2, nu=1.0, eps=1e-11, mu_inj=0.1)
    _say("fused vs unfused", n1, n2)
    out = []
    for variant in (0, 1, 2):
        o, g = pair(p, mg_variant=variant)
        c = smooth_c(o, seed=17)
        q = o.source("one_cell")
        g.set_concentration(c)
        g.set_source(q)
        its, rr = g.pressure_solve()
        assert rr <= 1e-11
        out.append((its, o.interior(g.get_pressure())))
        del capi
        from impec2d_amd import capi  # noqa: F401
    _say("  iterations: streaming %d, one kernel per pass %d, legs to the coarsest %d"
         % (out[0][0], out[1][0], out[2][0]))
    assert abs(out[0][0] - out[1][0]) <= 3, (out[0][0], out[1][0])
    assert abs(out[0][0] - out[2][0]) <= 1, (out[0][0], out[2][0])
    assert rel_err(out[0][1], out[1][1]) < 1e-9
    assert rel_err(out[0][1], out[2][1]) < 1e-9


@pytest.mark.parametrize("n1,n2", [(256, 256), (200, 120), (77, 45), (2048, 2048)])
def test_two_sweep_legs_match_the_unfused_cycle(gpu_lib, n1, n2):
    """mg_pre = mg_post = 2: the streaming legs with two red-black pairs each
    (wider strip halos, the bottom kernel with two pairs too) against the same
    legs run to the coarsest level, the one-kernel-per-pass V(2,2) cycle and
    the V(1,1) legs: all reach the tolerance with the same p, the two-pair
    cycles within one iteration of each other and no slower than V(1,1)."""
    p = params(N1=n1, N2=n2, nu=1.0, eps=1e-11, mu_inj=0.1)
    _say("two-sweep legs", n1, n2)
    out = []
    for variant, ns in ((0, 2), (2, 2), (1, 2), (0, 1)):
        o, g = pair(p, mg_variant=variant, mg_pre=ns, mg_post=ns)
        c = smooth_c(o, seed=17)
        g.set_concentration(c)
        g.set_source(o.source("one_cell"))
        its, rr = g.pressure_solve()
        assert rr <= 1e-11
        out.append((its, o.interior(g.get_pressure())))
    _say("  iterations: legs V(2,2) %d, legs to the coarsest %d, unfused V(2,2) %d, legs V(1,1) %d"
         % tuple(x[0] for x in out))
    assert abs(out[0][0] - out[2][0]) <= 1 and out[0][0] <= out[3][0], [x[0] for x in out]
    assert abs(out[0][0] - out[1][0]) <= 1
    for x in out[1:]:
        assert rel_err(out[0][1], x[1]) < 1e-9


@pytest.mark.parametrize("kw", [dict(nu=1.0), dict(nu=1.0, diffusion=1, alpha_abs=1, alpha_l=1e-3, alpha_t=1e-4,
                                                   d_m=1e-5)])
def test_steps_with_two_sweep_legs(gpu_lib, kw):
    """Ten IMPEC steps with the two-pair streaming legs on a grid large enough
    for them (96^2: two streaming levels above the bottom kernel) agree with
    the oracle to 1e-10 relative."""
    p = params(**dict(dict(N1=96, N2=96, eps=1e-13, para_eps=1e-13, mu_inj=0.1), **kw))
    _say("steps, two-sweep legs", kw)
    o, g = pair(p, mg_pre=2, mg_post=2)
    q = o.source("one_cell")
    c, pr, v = o.scalar(), o.scalar(), o.velocity()
    g.set_source(q)
    g.set_concentration(c)
    t_ref = t_gpu = 0.0
    for k in range(1, 11):
        st, dt_ref = o.step(c, pr, v, q, t_ref)
        assert st == 0
        t_ref += dt_ref
        t_gpu += g.step(t_gpu)
    cg, pg, vg = g.get_concentration(), g.get_pressure(), g.get_velocity()
    assert rel_err(o.interior(pg), o.interior(pr)) < 1e-10
    assert rel_err(vg, v) < 1e-10
    assert rel_err(o.interior(cg), o.interior(c)) < 1e-10
