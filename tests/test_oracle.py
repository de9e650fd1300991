"""Pins the CPU oracle (oracle/impec_oracle.c) to facts it did not compute:
closed forms, textbook reductions, published convergence orders of the
reference (examples/elliptic.eps, examples/hyperbolic.eps), invariants and
brute force.  CPU only."""
import math

import numpy as np
import pytest
from scipy.fft import dctn, idctn

from oracle import pyoracle as po
from tests.helpers import params, random_c, smooth_c


# ---------------------------------------------------------------------------
# pressure operator and solve
# ---------------------------------------------------------------------------
def dct_pinned_solution(N1, N2, dx, dy, nu, q):
    """Constant coefficients (M = 1, K = 1): the Neumann-folded nu-blend of the
    5-point and the 45-degree-rotated stencil is diagonalised by the 2-D
    DCT-II.  Returns the solution pinned to p(1,1) = 0 (q(1,1) = 0)."""
    tk = np.pi * np.arange(N1) / N1
    tl = np.pi * np.arange(N2) / N2
    ck, cl = np.cos(tk)[:, None], np.cos(tl)[None, :]
    lam_par = (2 * ck - 2) / dx ** 2 + (2 * cl - 2) / dy ** 2
    # rotated stencil for dx == dy: weight 1/(2 dx^2) on the four diagonal neighbours
    lam_rot = 4 * (ck * cl - 1) / (2 * dx * dy)
    lam = nu * lam_par + (1 - nu) * lam_rot
    b = -np.asarray(q).reshape(N1, N2)
    B = dctn(b, type=2, norm="ortho")
    lam[0, 0] = 1.0
    B[0, 0] = 0.0
    p = idctn(B / lam, type=2, norm="ortho")
    return p - p[1, 1]


@pytest.mark.parametrize("nu,N1,N2", [(1.0, 16, 16), (2 / 3.0, 16, 16), (0.0, 20, 20), (1.0, 24, 12),
                                      (2 / 3.0, 32, 32)])
def test_pressure_matches_dct_closed_form(nu, N1, N2):
    p = params(N1=N1, N2=N2, nu=nu, mu_res=1.0, mu_inj=1.0, eps=1e-14)
    o = po.Oracle(p)
    c = o.scalar()
    o.fill(c)
    q = o.source("one_cell")
    exact = dct_pinned_solution(N1, N2, o.dx, o.dy, nu, q)
    pr = o.scalar()
    assert o.pressure_solve(c, pr, q) == 0
    scale = np.abs(exact).max()
    assert np.abs(o.interior(pr) - exact).max() / scale < 1e-11
    if N1 * N2 <= 1024:
        pd = o.scalar()
        assert o.pressure_solve_dense(c, pd, q) == 0
        assert np.abs(o.interior(pd) - exact).max() / scale < 1e-12


def test_five_point_operator_is_textbook_harmonic_stencil():
    """nu = 1: interior rows are the 5-point operator with harmonic-mean face
    mobilities K/mu, lambda = 2 / (mu_l + mu_r)."""
    p = params(N1=8, N2=6, nu=1.0, mu_res=3.0, mu_inj=0.5)
    o = po.Oracle(p)
    c = random_c(o, seed=3)
    A = o.assemble(c).reshape(9, 8, 6)
    m4 = (3.0 / 0.5) ** 0.25
    mu = 3.0 / (1 - c + m4 * c) ** 4
    b = o.b
    for i in range(2, 6):
        for j in range(2, 4):
            lw = 2 / (mu[b + i - 1, b + j] + mu[b + i, b + j])
            le = 2 / (mu[b + i, b + j] + mu[b + i + 1, b + j])
            ls = 2 / (mu[b + i, b + j - 1] + mu[b + i, b + j])
            ln = 2 / (mu[b + i, b + j] + mu[b + i, b + j + 1])
            row = A[:, i, j]
            np.testing.assert_allclose(row[1], lw / o.dx ** 2, rtol=1e-14)
            np.testing.assert_allclose(row[5], le / o.dx ** 2, rtol=1e-14)
            np.testing.assert_allclose(row[3], ls / o.dy ** 2, rtol=1e-14)
            np.testing.assert_allclose(row[7], ln / o.dy ** 2, rtol=1e-14)
            np.testing.assert_allclose(row[0], -(lw + le) / o.dx ** 2 - (ls + ln) / o.dy ** 2, rtol=1e-14)
            assert row[2] == row[4] == row[6] == row[8] == 0


def test_rotated_operator_constant_coefficients():
    """nu = 0, dx == dy, K = mu = 1: weight 1/(2 dx^2) on the diagonal
    neighbours, zero on the axis neighbours, -2/dx^2 on the centre."""
    o = po.Oracle(params(N1=6, N2=6, nu=0.0, mu_res=1.0, mu_inj=1.0))
    c = o.scalar()
    A = o.assemble(c).reshape(9, 6, 6)
    row = A[:, 3, 3]
    h2 = o.dx ** 2
    np.testing.assert_allclose(row[[2, 4, 6, 8]], 1 / (2 * h2), rtol=1e-14)
    np.testing.assert_allclose(row[0], -2 / h2, rtol=1e-14)
    assert np.all(np.abs(row[[1, 3, 5, 7]]) <= 1e-12 / h2)


def test_pin_row_and_neumann_row_sums():
    o = po.Oracle(params(N1=10, N2=10, nu=2 / 3.0))
    c = random_c(o, seed=4)
    A = o.assemble(c).reshape(9, 10, 10)
    assert A[0, 1, 1] == 1.0 / (o.dx * o.dy) and np.all(A[1:, 1, 1] == 0)
    rows = A.sum(axis=0)
    rows[1, 1] = 0
    assert np.abs(rows).max() < 1e-9 * np.abs(A[0]).max()


# manufactured solutions of EllipticSolver::TestNeumann (EllipticSolver.cpp:319-470)
def _K(test, x, y):
    t = 2.0
    if test == 1:
        return np.ones_like(x), np.ones_like(x)
    if test == 2:
        return np.exp(-0.3 * x), np.exp(-0.3 * y)
    if test == 3:
        k = 2 + np.sin(t * x) * np.sin(y)
        return k, k
    return 2 + np.sin(t * x) * np.sin(y), 2 + np.sin(x) * np.sin(t * y)


def _q(test, x, y):
    t = 2.0
    if test == 1:
        return 2 * np.cos(x) * np.cos(y)
    if test == 2:
        a1, a2 = np.exp(-0.3 * x), np.exp(-0.3 * y)
        a1x, a2y = -0.3 * a1, -0.3 * a2
    elif test == 3:
        a1 = a2 = 2 + np.sin(t * x) * np.sin(y)
        a1x, a2y = t * np.cos(t * x) * np.sin(y), np.cos(y) * np.sin(t * x)
    else:
        a1, a2 = 2 + np.sin(t * x) * np.sin(y), 2 + np.sin(x) * np.sin(t * y)
        a1x, a2y = t * np.cos(t * x) * np.sin(y), t * np.cos(t * y) * np.sin(x)
    return (a1 + a2) * np.cos(x) * np.cos(y) + a1x * np.sin(x) * np.cos(y) + a2y * np.cos(x) * np.sin(y)


def manufactured_errors(test, nu, N):
    L = 2 * math.pi
    p = params(N1=N, N2=N, nu=nu, x_ub=L, y_ub=L, mu_res=1.0, mu_inj=1.0, avgproc=1, eps=1e-13)
    dx = L / N
    b = 2
    xc = (np.arange(-b, N + b) + 0.5) * dx
    X, Y = np.meshgrid(xc, xc, indexing="ij")
    K11, K22 = _K(test, X, Y)
    o = po.Oracle(p, np.ascontiguousarray(K11), np.ascontiguousarray(K22))
    xi = (np.arange(N) + 0.5) * dx
    Xi, Yi = np.meshgrid(xi, xi, indexing="ij")
    q = _q(test, Xi, Yi).reshape(-1)
    pex = np.cos(Xi) * np.cos(Yi)
    o.set_kappa(-dx * dx * q[1 * N + 1] / pex[1, 1])   # MGSolver.cpp:528-548
    c, pr, v = o.scalar(), o.scalar(), o.velocity()
    o.fill(c)
    assert o.pressure_solve(c, pr, q) == 0
    o.velocities(pr, v, c)
    ep = np.abs(o.interior(pr) - pex).sum() * dx * dx
    xu = np.arange(N + 1) * dx
    Xu, Yu = np.meshgrid(xu, xi, indexing="ij")
    uex = _K(test, Xu, Yu)[0] * np.sin(Xu) * np.cos(Yu)
    eu = np.abs(v[b:b + N + 1, b:b + N, 0] - uex).sum() * dx * dx
    return dx, ep, eu


@pytest.mark.parametrize("test,nu", [(1, 1.0), (1, 0.0), (1, 2 / 3.0), (2, 1.0), (3, 2 / 3.0), (4, 1.0)])
def test_manufactured_convergence_order(test, nu):
    """Second order in pressure and velocity, as published for the unrotated
    (2.00/2.00), rotated (2.00/1.99) and blended (2.00/1.99) stencils."""
    res = [manufactured_errors(test, nu, N) for N in (25, 50, 100)]
    dx = np.log([r[0] for r in res])
    kp = np.polyfit(dx, np.log([r[1] for r in res]), 1)[0]
    ku = np.polyfit(dx, np.log([r[2] for r in res]), 1)[0]
    assert 1.85 < kp < 2.15, kp
    assert 1.8 < ku < 2.2, ku


def test_divergence_and_noflow_after_solve():
    for nu in (1.0, 2 / 3.0):
        o = po.Oracle(params(N1=40, N2=40, nu=nu, eps=1e-13))
        c = smooth_c(o, seed=5)
        q = o.source()
        pr, v = o.scalar(), o.velocity()
        assert o.pressure_solve(c, pr, q) == 0
        o.velocities(pr, v, c)
        b = o.b
        div = ((v[b:-b - 1, b + 1:-b, 1] - v[b:-b - 1, b:-b - 1, 1]) / o.dy
               + (v[b + 1:-b, b:-b - 1, 0] - v[b:-b - 1, b:-b - 1, 0]) / o.dx)
        err = np.abs(div - q.reshape(40, 40))
        err[1, 1] = 0
        assert err.max() < 1e-7 * np.abs(q).max()
        nf = o.noflow_error(v)
        assert (nf == 0.0) if nu == 1.0 else (nf < 1e-10 * np.abs(v).max())


# ---------------------------------------------------------------------------
# time step, sources, ghosts
# ---------------------------------------------------------------------------
def test_cfl_closed_form():
    o = po.Oracle(params(N1=10, N2=20, cfl=0.5))
    v = o.velocity()
    v[..., 0] = 3.0
    v[..., 1] = -1.5
    assert o.timestep(v) == 0.5 * (min(o.dx, o.dy) / 3.0)


def test_tr_closed_form():
    o = po.Oracle(params(N1=8, N2=8, inj_fluid_conc=0.7))
    q = o.source("one_cell")
    c = o.scalar()
    o.interior(c)[:] = 0.4
    dt = 1e-4
    o.tr(c, q, dt)
    h = 0.5 * dt * q[-1]
    assert abs(o.interior(c)[-1, -1] - 0.4 * (1 + h) / (1 - h)) < 1e-15
    assert abs(o.interior(c)[0, 0] - (0.4 + dt * 0.7 * q[0])) < 1e-12
    assert np.all(o.interior(c).reshape(-1)[1:-1] == 0.4)


def test_neumann_ghosts_mirror():
    o = po.Oracle(params(N1=7, N2=5))
    c = o.scalar()
    o.interior(c)[:] = np.random.default_rng(0).random((7, 5))
    o.fill_neumann(c)
    b = o.b
    for k in range(b):
        assert np.array_equal(c[b - 1 - k, b:-b], c[b + k, b:-b])
        assert np.array_equal(c[b:-b, b - 1 - k], c[b:-b, b + k])
    assert c[0, 0] == c[b + 1, b + 1] and c[-1, -1] == c[-b - 2, -b - 2]


def test_velocity_ghosts_parity():
    """Normal components are odd about the walls, tangential ones even."""
    o = po.Oracle(params(N1=6, N2=5))
    v = np.random.default_rng(1).random(o.velocity().shape)
    o.fill_velocity(v)
    b, N1, N2 = o.b, 6, 5
    for k in (1, 2):
        assert np.array_equal(v[b - k, b:b + N2, 0], -v[b + k, b:b + N2, 0])
        assert np.array_equal(v[b:b + N1, b - k, 1], -v[b:b + N1, b + k, 1])
        assert np.array_equal(v[b + N1 - 1 + k, b:b + N2 + 1, 1], v[b + N1 - k, b:b + N2 + 1, 1])


def test_circle_well_area_closed_form():
    """Exact cell area fractions of a quarter disc add up to pi r^2 / 4."""
    for r, n in [(0.3, 16), (0.05, 64), (0.17, 37)]:
        h = 1.0 / n
        tot = sum(po.exact_vol_frac_circle(i * h, j * h, r, h, h) for i in range(n) for j in range(n))
        assert abs(tot * h * h - math.pi * r * r / 4) < 1e-13


def test_circle_sources_balance():
    for stype in (0, 1):
        o = po.Oracle(params(N1=40, N2=40, stype=stype, well_rc=0.1))
        q = o.source("circle").reshape(40, 40)
        assert abs(q.sum()) < 1e-9 * np.abs(q).max()
        inj = q[q > 0].sum() * o.dx * o.dy
        assert abs(inj - (2.0 if stype == 0 else 1.0)) < 1e-12


# ---------------------------------------------------------------------------
# transport
# ---------------------------------------------------------------------------
def minmod_muscl_x(c, u, dt, dx, b):
    """Independent MUSCL-Hancock x-sweep with the minmod slope (Toro), on the
    reference's index ranges; porosity 1."""
    cc = c.copy()
    N1 = c.shape[0] - 2 * b

    def mm(a, d):
        return np.where(a * d > 0, np.sign(a) * np.minimum(np.abs(a), np.abs(d)), 0.0)
    cm, c0, cp = c[b - 2:b + N1, :], c[b - 1:b + N1 + 1, :], c[b:b + N1 + 2, :]
    d = mm(c0 - cm, cp - c0)
    UL, UR = c0 - 0.5 * d, c0 + 0.5 * d
    ul, ur = u[b - 1:b + N1 + 1, :c.shape[1], 0], u[b:b + N1 + 2, :c.shape[1], 0]
    df = 0.5 * dt / dx * (ul * UL - ur * UR)
    left, right = UR + df, UL + df            # at faces k+1 and k of cell k (k = -1 .. N1)
    uf = u[b:b + N1 + 1, :c.shape[1], 0]
    F = np.where(uf > 0, uf * left[:-1], uf * right[1:])
    cc[b:b + N1, :] = c[b:b + N1, :] + dt / dx * (F[:-1] - F[1:])
    return cc


def test_muscl_ratio_limiter_equals_minmod():
    o = po.Oracle(params(N1=30, N2=20))
    c = smooth_c(o, seed=6)
    rng = np.random.default_rng(7)
    v = rng.standard_normal(o.velocity().shape)
    dt = 0.1 / 30
    ref = minmod_muscl_x(c, v, dt, o.dx, o.b)
    o.advect_x(c, v, dt)
    assert np.abs(o.interior(c) - o.interior(ref)).max() < 1e-13


def sine_profile(x, y, dx, dy):
    # Sinewave (Util.cpp:88-92): cell average of a product sine wave
    return 1 + 1 / (dx * dy * 4 * np.pi ** 2) * ((np.cos(2 * np.pi * (x + dx / 2)) - np.cos(2 * np.pi * (x - dx / 2)))
                                                 * (np.cos(2 * np.pi * (y + dy / 2)) - np.cos(2 * np.pi * (y - dy / 2))))


def advection_l1(N, hyper, cfl=0.9):
    """TestAdvection/Advection (Impec2DSolver.cpp:773-990): periodic unit
    square, u = (-1, 1), one period, L1 error against the initial profile."""
    o = po.Oracle(params(N1=N, N2=N, is_test=1, hyper=hyper, cfl=cfl, dim_split=4))
    x = (np.arange(N) + 0.5) / N
    X, Y = np.meshgrid(x, x, indexing="ij")
    c = o.scalar()
    ex = sine_profile(X, Y, 1.0 / N, 1.0 / N)
    o.interior(c)[:] = ex
    v = o.velocity()
    v[..., 0] = -1.0
    v[..., 1] = 1.0
    dtmax = o.L.or_max_stable_advection_timestep(o.ctx, v)
    t = 0.0
    while t < 1.0:
        dt = 1.0 - t if t + dtmax >= 1.0 else dtmax
        o.advect_split(c, v, dt)
        t += dt
    return np.abs(ex - o.interior(c)).sum() / (N * N), c


def test_muscl_and_upwind_convergence_orders():
    """Published: MUSCL 1.85, upwind 1.00 (examples/hyperbolic.eps:882, 945)."""
    Ns = (50, 100, 200)
    for hyper, lo, hi in ((0, 1.7, 2.0), (1, 0.85, 1.1)):
        errs = [advection_l1(N, hyper)[0] for N in Ns]
        k = -np.polyfit(np.log(Ns), np.log(errs), 1)[0]
        assert lo < k < hi, (hyper, k)


def test_transport_conserves_and_preserves_constants():
    o = po.Oracle(params(N1=20, N2=20, is_test=1))
    v = o.velocity()
    rng = np.random.default_rng(8)
    v[..., 0] = rng.uniform(0.5, 1.0)
    v[..., 1] = rng.uniform(-1.0, -0.5)
    c = o.scalar()
    o.interior(c)[:] = 0.37
    o.advect_split(c, v, 0.01)
    assert np.all(o.interior(c) == 0.37)
    c = smooth_c(o, seed=9)
    m0 = o.interior(c).sum()
    o.advect_split(c, v, 0.01)
    assert abs(o.interior(c).sum() - m0) < 1e-13 * m0


def test_parallel_splitting_is_transpose_symmetric():
    """Exchanging x and y maps the XY/YX-averaged update onto itself."""
    o = po.Oracle(params(N1=16, N2=16, dim_split=4))
    c = random_c(o, seed=10)
    rng = np.random.default_rng(11)
    v = rng.standard_normal(o.velocity().shape)
    vt = np.stack([v[..., 1].T, v[..., 0].T], axis=-1).copy()
    ct = c.T.copy()
    o.advect_split(c, v, 0.004)
    o.advect_split(ct, vt, 0.004)
    assert np.array_equal(o.interior(c), o.interior(ct).T)


# ---------------------------------------------------------------------------
# dispersion
# ---------------------------------------------------------------------------
def test_cn_solve_matches_dense_and_keeps_constants():
    o = po.Oracle(params(N1=12, N2=12, diffusion=1, alpha_l=0.5, alpha_t=0.2, para_eps=1e-14))
    rng = np.random.default_rng(12)
    v = rng.standard_normal(o.velocity().shape) + 0.3
    o.fill_velocity(v)
    c = smooth_c(o, seed=13)
    dt = 0.004
    x = po.dense_solve9(12, 12, o.parabolic_matrix(v, dt), o.parabolic_rhs(c, v, dt))
    cc = c.copy()
    assert o.parabolic_solve(cc, v, dt) == 0
    assert np.abs(o.interior(cc).reshape(-1) - x).max() < 1e-12
    k = o.scalar()
    k[:] = 0.61
    assert o.parabolic_solve(k, v, dt) == 0
    assert np.abs(o.interior(k) - 0.61).max() < 1e-12


def test_cn_is_first_order_dispersion():
    """Small dt: (c_new - c)/dt approaches div(D grad c) on a smooth field:
    halving dt leaves the increment rate unchanged to O(dt)."""
    o = po.Oracle(params(N1=24, N2=24, diffusion=1, alpha_abs=1, alpha_l=2e-3, alpha_t=1e-3, d_m=1e-3,
                         para_eps=1e-14))
    v = o.velocity()
    v[..., 0] = 1.0
    v[..., 1] = 0.5
    o.fill_velocity(v)
    c = smooth_c(o, seed=14)
    rates = []
    for dt in (1e-4, 5e-5):
        cc = c.copy()
        o.parabolic_solve(cc, v, dt)
        rates.append((o.interior(cc) - o.interior(c)) / dt)
    assert np.abs(rates[0] - rates[1]).max() < 1e-2 * np.abs(rates[1]).max()


# ---------------------------------------------------------------------------
# whole step invariants
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("kw", [dict(), dict(stype=0, x_ub=2 ** 0.5, y_ub=2 ** 0.5, nu=2 / 3.0),
                                dict(diffusion=1, alpha_abs=1, alpha_l=1e-3, alpha_t=1e-4, d_m=1e-5)])
def test_mass_balance_over_steps(kw):
    o = po.Oracle(params(**dict(dict(N1=32, N2=32, eps=1e-12, para_eps=1e-12), **kw)))
    c, pr, v, q = o.scalar(), o.scalar(), o.velocity(), o.source()
    t = 0.0
    for _ in range(15):
        st, dt = o.step(c, pr, v, q, t)
        assert st == 0
        t += dt
    assert o.mass_error(c, t) < 1e-10
    assert o.div_error(v, q) < 1e-6


def test_diagonal_case_is_transpose_symmetric_to_tolerance():
    o = po.Oracle(params(N1=24, N2=24, eps=1e-13))
    c, pr, v, q = o.scalar(), o.scalar(), o.velocity(), o.source()
    t = 0.0
    for _ in range(8):
        st, dt = o.step(c, pr, v, q, t)
        t += dt
    ci = o.interior(c)
    assert np.abs(ci - ci.T).max() < 1e-10 * np.abs(ci).max()
