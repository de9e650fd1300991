"""Ensembles (impec_ensemble_*): members of different sizes, grids, mobility
ratios and lognormal permeabilities advanced together by batched launches and
graph replays.  Each member must equal, bit for bit, the same problem stepped
alone on its own context with impec_step (which the other GPU tests hold to
the oracle); graph replays and eager launches must agree; chunk heights
picked by the autotuner keep that equality."""
import numpy as np
import pytest

from helpers import gpu_params, pair, params, smooth_c
from impec2d_amd import fields
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

R2 = 2 ** 0.5
# (N1, N2, grid, mobility ratio M, sigma of ln K); grid "diag": injector and
# producer at opposite corners of the unit square, "par": the rotated square
# with four corner wells
MEMBERS = [
    (32, 32, "diag", 10.0, 0.5),
    (48, 40, "par", 30.0, 1.5),
    (64, 64, "diag", 50.0, 2.0),
    (130, 70, "par", 2.0, 1.0),
    (200, 256, "diag", 20.0, 1.0),
]


def _say(*a):
    print(*a, flush=True)


def _member_setup(k, spec, **extra):
    n1, n2, grid, M, sigma = spec
    kw = dict(N1=n1, N2=n2, nu=1.0, eps=1e-10, mu_inj=1.0 / M)
    if grid == "par":
        kw.update(stype=0, x_ub=R2, y_ub=R2)
    p = params(**kw)
    K = fields.lognormal_permeability(n1, n2, sigma=sigma, corr_len=0.05, seed=100 + k,
                                      x_extent=kw.get("x_ub", 1.0), y_extent=kw.get("y_ub", 1.0))
    return p, K


def _run_single(p, K, steps, **extra):
    o, g = pair(p, K, None, **extra)
    q = o.source("one_cell")
    c0 = smooth_c(o, seed=3)
    g.set_source(q)
    g.set_concentration(c0)
    out = []
    t = 0.0
    for _ in range(steps):
        dt = g.step(t)
        t += dt
        out.append((dt, g.stats().last_its, g.get_concentration(), g.get_pressure(), g.get_velocity()))
    g.close()
    return q, c0, out


def _make_ensemble(specs, rows=None, **extra):
    from impec2d_amd import capi
    plist, Ks = [], []
    for k, spec in enumerate(specs):
        p, K = _member_setup(k, spec)
        gp = gpu_params(p, **extra)
        if rows is not None:
            gp.mg_rows = int(rows[k])
        plist.append(gp)
        Ks.append((p, K))
    ens = capi.Ensemble(plist)
    for k, (p, K) in enumerate(Ks):
        o = po.Oracle(p, K, None)   # host-side source and initial state
        m = ens.members[k]
        m.set_permeability(K, None)
        m.set_source(o.source("one_cell"))
        m.set_concentration(smooth_c(o, seed=3))
    return ens, Ks


def _ensemble_steps(ens, steps):
    res = []
    for _ in range(steps):
        dt, its = ens.step()
        res.append((dt.copy(), its.copy(), [(m.get_concentration(), m.get_pressure(), m.get_velocity())
                                             for m in ens.members]))
    return res


@pytest.mark.parametrize("precision,guess", [(0, 0), (1, 1)])
def test_members_equal_single_contexts_bitwise(gpu_lib, precision, guess):
    steps = 3
    _say("ensemble vs single contexts: precision %d guess %d" % (precision, guess))
    ens, Ks = _make_ensemble(MEMBERS, precision=precision, initial_guess=guess)
    res = _ensemble_steps(ens, steps)
    ens.close()
    for k, (p, K) in enumerate(Ks):
        _, _, single = _run_single(p, K, steps, precision=precision, initial_guess=guess)
        for s in range(steps):
            dt_e, its_e, fields_e = res[s][0][k], res[s][1][k], res[s][2][k]
            dt_s, its_s, c_s, p_s, v_s = single[s]
            c_e, p_e, v_e = fields_e
            _say("  member %d step %d: its %d/%d dt %.6e" % (k, s, its_e, its_s, dt_e))
            assert its_e == its_s
            assert dt_e == dt_s
            assert np.array_equal(p_e, p_s)
            assert np.array_equal(v_e, v_s)
            assert np.array_equal(c_e, c_s)


def test_large_member_equal_single_context_bitwise(gpu_lib):
    """A 2048^2 member leaves more partials than one workgroup combines quickly
    (16384 after the CG update): the single context combines them in two
    launches, the ensemble in one workgroup per member, and the two must agree
    bit for bit."""
    steps = 2
    specs = [(2048, 2048, "par", 10.0, 1.0), (64, 48, "diag", 20.0, 1.0)]
    _say("ensemble with a 2048^2 member vs single contexts")
    ens, Ks = _make_ensemble(specs)
    res = _ensemble_steps(ens, steps)
    ens.close()
    for k, (p, K) in enumerate(Ks):
        _, _, single = _run_single(p, K, steps)
        for s in range(steps):
            dt_e, its_e, (c_e, p_e, v_e) = res[s][0][k], res[s][1][k], res[s][2][k]
            dt_s, its_s, c_s, p_s, v_s = single[s]
            _say("  member %d step %d: its %d/%d" % (k, s, its_e, its_s))
            assert its_e == its_s and dt_e == dt_s
            assert np.array_equal(p_e, p_s)
            assert np.array_equal(v_e, v_s)
            assert np.array_equal(c_e, c_s)


def test_graph_replays_match_eager_launches(gpu_lib):
    """The same ensemble stepped with eager launches, with graphs of one CG
    iteration per replay and with graphs of four: identical fields after
    every step (each step replays the iteration graph several times, with the
    host's synchronisation and reads in between)."""
    steps = 3
    runs = {}
    for mode in [("eager", 0, 0), ("graph1", 1, 1), ("graph4", 1, 4)]:
        _say("ensemble launch mode", mode[0])
        ens, _ = _make_ensemble(MEMBERS[:4], precision=1)
        ens.set_graphs(mode[1], mode[2])
        runs[mode[0]] = _ensemble_steps(ens, steps)
        ens.close()
    for name in ("graph1", "graph4"):
        for s in range(steps):
            a, b = runs["eager"][s], runs[name][s]
            assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
            for fa, fb in zip(a[2], b[2]):
                for x, y in zip(fa, fb):
                    assert np.array_equal(x, y)


def test_autotuned_rows_keep_member_equality(gpu_lib):
    specs = MEMBERS[:3] + [(512, 512, "diag", 20.0, 1.0)]
    ens, Ks = _make_ensemble(specs)
    rows = ens.autotune(reps=2)
    _say("autotuned finest-level chunk rows", rows.tolist())
    assert all(r in (8, 16, 32, 64) for r in rows)
    for k in range(len(specs)):
        for q in range(len(specs)):
            if specs[q][:2] == specs[k][:2]:
                assert rows[q] == rows[k]
    res = _ensemble_steps(ens, 2)
    ens.close()
    for k, (p, K) in enumerate(Ks):
        _, _, single = _run_single(p, K, 2, mg_rows=int(rows[k]))
        for s in range(2):
            assert res[s][1][k] == single[s][1]
            c_e, p_e, v_e = res[s][2][k]
            assert np.array_equal(p_e, single[s][3]) and np.array_equal(c_e, single[s][2])
