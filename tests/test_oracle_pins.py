"""Pins of the oracle where the reference's own tests leave it open: the
Crank-Nicolson dispersion assembly (MGSolver.cpp:846-948, 997-1370) and the
sequencing of the splitting modes (HyperbolicSplitSolver.cpp:51-106).

Each pin is a fact the mathematics fixes, computed here without the oracle's
formulas:
  * for a uniform velocity the dispersion tensor D = phi d_m I +
    phi |u| (alpha_l E + alpha_t (I - E)), E = u u^T / |u|^2, is constant, and
    every interior row of the CN matrix is I - (dt/2)(D11 dxx + D22 dyy +
    2 D12 dxy) of the textbook central differences (D from a 2x2 matrix
    product here, not from the oracle's component formulas);
  * the explicit half (the right-hand side) is exact on quadratics:
    rhs - c = dt D11, dt D22, dt D12 for c = x^2, y^2, xy;
  * for a velocity field linear in x and y the discrete operator is a
    second-order consistent approximation of div(D grad c), measured against
    a finite-difference evaluation of the continuous operator;
  * every splitting mode equals the composition, in the mode's order and with
    the mode's step fractions, of an independent numpy MUSCL-Hancock sweep.

Every pin is also shown to fail on a mistake planted in the oracle
(or_set_mutation): alpha_l/alpha_t exchanged, porosity dropped, d_m in the
cross term, wrong four-point-average indices, a cross-term sign of the
right-hand side, and wrong sweep orders or step fractions."""
import numpy as np
import pytest

from oracle import pyoracle as po
from tests.helpers import params, random_c
from tests.test_oracle import minmod_muscl_x


# ---------------------------------------------------------------------------
# dispersion tensor from the physics (a 2x2 matrix product)
# ---------------------------------------------------------------------------
def dispersion_tensor(U, V, phi, al, at, dm):
    u = np.array([U, V], dtype=float)
    s = np.linalg.norm(u)
    E = np.outer(u, u) / s ** 2
    I = np.eye(2)
    return phi * dm * I + phi * s * (al * E + at * (I - E))


UNIFORM = dict(N1=12, N2=9, x_ub=1.0, y_ub=0.8, diffusion=1, alpha_l=0.02, alpha_t=0.005, d_m=1e-3,
               porosity=0.7)
U0, V0 = 1.3, -0.7


def _uniform_case(alpha_abs):
    o = po.Oracle(params(alpha_abs=alpha_abs, **UNIFORM))
    v = o.velocity()
    v[..., 0] = U0
    v[..., 1] = V0
    scale = 1.0 if alpha_abs else max(o.dx, o.dy)   # relative alpha: times max(dx, dy)
    D = dispersion_tensor(U0, V0, UNIFORM["porosity"], UNIFORM["alpha_l"] * scale, UNIFORM["alpha_t"] * scale,
                          UNIFORM["d_m"])
    return o, v, D


def uniform_matrix_error(alpha_abs):
    """max relative deviation of the interior rows from I - dt/2 L_D."""
    o, v, D = _uniform_case(alpha_abs)
    N1, N2, dx, dy = o.N1, o.N2, o.dx, o.dy
    dt = 3e-3
    A = o.parabolic_matrix(v, dt)
    x = np.random.default_rng(1).standard_normal((N1, N2))
    y = po.matvec9(N1, N2, A, x.ravel()).reshape(N1, N2)
    xc = x[1:-1, 1:-1]
    dxx = (x[2:, 1:-1] - 2 * xc + x[:-2, 1:-1]) / dx ** 2
    dyy = (x[1:-1, 2:] - 2 * xc + x[1:-1, :-2]) / dy ** 2
    dxy = (x[2:, 2:] - x[2:, :-2] - x[:-2, 2:] + x[:-2, :-2]) / (4 * dx * dy)
    ref = xc - dt / 2 * (D[0, 0] * dxx + D[1, 1] * dyy + 2 * D[0, 1] * dxy)
    return np.abs(y[1:-1, 1:-1] - ref).max() / np.abs(ref).max()


def rhs_quadratic_error(alpha_abs):
    """max relative deviation of rhs - c from dt*(D11, D22, D12) for c = x^2,
    y^2, xy at the cells whose stencil stays inside the grid."""
    o, v, D = _uniform_case(alpha_abs)
    b, dx, dy = o.b, o.dx, o.dy
    dt = 0.5
    x = (np.arange(-b, o.N1 + b) + 0.5) * dx
    y = (np.arange(-b, o.N2 + b) + 0.5) * dy
    X, Y = np.meshgrid(x, y, indexing="ij")
    worst = 0.0
    for f, expect in ((X ** 2, dt * D[0, 0]), (Y ** 2, dt * D[1, 1]), (X * Y, dt * D[0, 1])):
        c = np.ascontiguousarray(f)
        r = o.parabolic_rhs(c, v, dt).reshape(o.N1, o.N2)
        got = (r - o.interior(c))[1:-1, 1:-1]
        worst = max(worst, np.abs(got - expect).max() / abs(expect))
    return worst


@pytest.mark.parametrize("alpha_abs", [0, 1])
def test_cn_matrix_uniform_velocity_closed_form(alpha_abs):
    assert uniform_matrix_error(alpha_abs) < 1e-13


@pytest.mark.parametrize("alpha_abs", [0, 1])
def test_cn_rhs_exact_on_quadratics(alpha_abs):
    assert rhs_quadratic_error(alpha_abs) < 1e-10


# ---------------------------------------------------------------------------
# second-order consistency with a velocity linear in x and y
# ---------------------------------------------------------------------------
PHI, AL, AT, DM = 0.8, 0.05, 0.01, 2e-3


def _U(X, Y):
    return 1.0 + 0.6 * X - 0.4 * Y


def _V(X, Y):
    return 0.4 - 0.5 * X + 0.7 * Y


def _c(X, Y):
    return np.sin(1.7 * X + 0.3) * np.cos(1.3 * Y - 0.2) + 0.5 * X * Y


def continuous_operator(X, Y, h=1e-4, e=1e-5):
    """div(D grad c) of the continuous fields, by fourth-order differences of
    the flux D grad c (itself from central differences of c)."""
    def flux(x, y):
        cx = (_c(x + e, y) - _c(x - e, y)) / (2 * e)
        cy = (_c(x, y + e) - _c(x, y - e)) / (2 * e)
        U, V = _U(x, y), _V(x, y)
        s = np.sqrt(U * U + V * V)
        d11 = PHI * DM + PHI * (AL * U * U + AT * V * V) / s
        d22 = PHI * DM + PHI * (AL * V * V + AT * U * U) / s
        d12 = PHI * (AL - AT) * U * V / s
        return d11 * cx + d12 * cy, d12 * cx + d22 * cy

    def d4(f, x, y, ax):
        sx, sy = (h, 0.0) if ax == 0 else (0.0, h)
        return (-f(x + 2 * sx, y + 2 * sy) + 8 * f(x + sx, y + sy) - 8 * f(x - sx, y - sy)
                + f(x - 2 * sx, y - 2 * sy)) / (12 * h)
    return d4(lambda a, b: flux(a, b)[0], X, Y, 0) + d4(lambda a, b: flux(a, b)[1], X, Y, 1)


def consistency_error(N, use_rhs):
    o = po.Oracle(params(N1=N, N2=N, x_ub=1.0, y_ub=1.0, diffusion=1, alpha_abs=1, alpha_l=AL, alpha_t=AT, d_m=DM,
                         porosity=PHI))
    b, h = o.b, o.dx
    v = o.velocity()
    ii = np.arange(-b, N + 1 + b)[:, None]
    jj = np.arange(-b, N + 1 + b)[None, :]
    v[..., 0] = _U(ii * h, (jj + 0.5) * h)      # x-face (i-1/2, j) at (i h, (j+1/2) h)
    v[..., 1] = _V((ii + 0.5) * h, jj * h)      # y-face (i, j-1/2) at ((i+1/2) h, j h)
    dt = 1e-3
    xc = (np.arange(N) + 0.5) * h
    X, Y = np.meshgrid(xc, xc, indexing="ij")
    xg = (np.arange(-b, N + b) + 0.5) * h
    c = np.ascontiguousarray(_c(*np.meshgrid(xg, xg, indexing="ij")))
    if use_rhs:
        Lh = (o.parabolic_rhs(c, v, dt).reshape(N, N) - _c(X, Y)) / (dt / 2)
    else:
        A = o.parabolic_matrix(v, dt)
        Lh = (_c(X, Y) - po.matvec9(N, N, A, _c(X, Y).ravel()).reshape(N, N)) / (dt / 2)
    sl = (slice(2, N - 2), slice(2, N - 2))
    return np.abs(Lh[sl] - continuous_operator(X, Y)[sl]).max()


def consistency_order(use_rhs):
    e = [consistency_error(N, use_rhs) for N in (32, 64, 128)]
    return np.log2(e[1] / e[2]), e


@pytest.mark.parametrize("use_rhs", [False, True])
def test_cn_operator_second_order_consistent(use_rhs):
    order, e = consistency_order(use_rhs)
    assert order > 1.85, (order, e)
    assert e[2] < 2e-5, e


# ---------------------------------------------------------------------------
# splitting modes as compositions of independent sweeps
# ---------------------------------------------------------------------------
def sweep_x(c, v, dt, h, b):
    return minmod_muscl_x(c, v, dt, h, b)


def sweep_y(c, v, dt, h, b):
    vt = np.ascontiguousarray(np.stack([v[..., 1].T, v[..., 0].T], axis=-1))
    return np.ascontiguousarray(minmod_muscl_x(np.ascontiguousarray(c.T), vt, dt, h, b).T)


def composed(mode, c, v, dt, h, b, first_x):
    """AdvectSplit's sweep order per mode (HyperbolicSplitSolver.cpp:51-106),
    from the ghost-filled state c; returns the new state and whether the next
    alternating step starts with x."""
    X = lambda a, t: sweep_x(a, v, t, h, b)   # noqa: E731
    Y = lambda a, t: sweep_y(a, v, t, h, b)   # noqa: E731
    if mode == 0:
        return (Y(X(c, dt), dt), False) if first_x else (X(Y(c, dt), dt), True)
    if mode == 1:
        return Y(X(c, dt), dt), first_x
    if mode == 2:
        return X(Y(c, dt), dt), first_x
    if mode == 3:
        return X(Y(X(c, dt / 2), dt), dt / 2), first_x
    return 0.5 * (Y(X(c, dt), dt) + X(Y(c, dt), dt)), first_x


def split_mode_error(mode, steps=2):
    """max deviation over the whole ghosted array after `steps` calls."""
    N = 20
    o = po.Oracle(params(N1=N, N2=N, dim_split=mode))
    c = random_c(o, seed=30 + mode)
    rng = np.random.default_rng(40 + mode)
    v = rng.standard_normal(o.velocity().shape)
    dt = 0.1 / N
    ref = c.copy()
    first_x = True                     # prev_first_sweep starts at Y_SWEEP (HyperbolicSolver.cpp:84)
    worst = 0.0
    for _ in range(steps):
        o.fill(ref)                    # AdvectSplit begins with FillBoundaries
        ref, first_x = composed(mode, ref, v, dt, o.dx, o.b, first_x)
        o.advect_split(c, v, dt)
        worst = max(worst, np.abs(c - ref).max())
        ref = c.copy()                 # the next call starts from the same state
    return worst


@pytest.mark.parametrize("mode", [0, 1, 2, 3, 4])
def test_split_modes_compose_independent_sweeps(mode):
    assert split_mode_error(mode) < 1e-13


def test_split_modes_differ():
    """The compositions are different operators on this data, so the pins
    above distinguish the orders."""
    N = 20
    o = po.Oracle(params(N1=N, N2=N))
    c = random_c(o, seed=31)
    o.fill(c)
    v = np.random.default_rng(41).standard_normal(o.velocity().shape)
    outs = [composed(m, c, v, 0.1 / N, o.dx, o.b, True)[0] for m in range(1, 5)]
    for a in range(4):
        for b in range(a + 1, 4):
            assert np.abs(outs[a] - outs[b]).max() > 1e-6


# ---------------------------------------------------------------------------
# each pin fails on a planted mistake
# ---------------------------------------------------------------------------
PLANTED = [
    ("cn_swap_alpha", lambda: uniform_matrix_error(0) < 1e-13),
    ("cn_swap_alpha", lambda: rhs_quadratic_error(1) < 1e-10),
    ("cn_drop_phi", lambda: uniform_matrix_error(1) < 1e-13),
    ("cn_drop_phi", lambda: rhs_quadratic_error(0) < 1e-10),
    ("cn_dm_cross", lambda: uniform_matrix_error(1) < 1e-13),
    ("cn_dm_cross", lambda: rhs_quadratic_error(1) < 1e-10),
    ("cn_rhs_cross", lambda: rhs_quadratic_error(0) < 1e-10),
    ("cn_avg_f", lambda: consistency_order(False)[0] > 1.85),
    ("cn_avg_f", lambda: consistency_order(True)[0] > 1.85),
    ("cn_avg_g", lambda: consistency_order(False)[0] > 1.85),
    ("cn_avg_g", lambda: consistency_order(True)[0] > 1.85),
    ("split_xfirst", lambda: split_mode_error(1) < 1e-13),
    ("split_yfirst", lambda: split_mode_error(2) < 1e-13),
    ("split_strang", lambda: split_mode_error(3) < 1e-13),
    ("split_alt", lambda: split_mode_error(0) < 1e-13),
    ("split_parallel", lambda: split_mode_error(4) < 1e-13),
]


@pytest.mark.parametrize("k", range(len(PLANTED)), ids=["%s-%d" % (m, k) for k, (m, _) in enumerate(PLANTED)])
def test_pins_fail_on_planted_mistake(k):
    name, pin_holds = PLANTED[k]
    assert pin_holds()                 # the faithful oracle passes
    with po.mutated(name):
        assert not pin_holds(), name   # the planted mistake is caught
