"""The solvers' lookahead: while the residual is far above the target, work
of the next step of the iteration is queued before the host reads the
residual that decides convergence.  Dispersion: an iteration's p-launch; when
the iteration turns out to end the solve, the p-launch stands in for the final
x update (the same expression on the same operands).  Pressure: the next
V-cycle's first leg, which reads only the residual.  The fields must not
depend on the lookahead factor: never (negative), the default (0: 1e6), and 1,
which looks ahead in every iteration that starts above the target, the
converging one included."""
import numpy as np
import pytest

from helpers import pair, params, rel_err

pytestmark = pytest.mark.gpu


def _run(ahead, steps=3):
    p = params(N1=48, N2=40, eps=1e-12, para_eps=1e-10, mu_inj=0.1, diffusion=1, alpha_abs=1,
               alpha_l=1e-2, alpha_t=1e-3, d_m=1e-4)
    o, g = pair(p, cn_ahead=ahead)
    q = o.source("one_cell")
    g.set_source(q)
    g.set_concentration(o.scalar())
    t = 0.0
    its = []
    for _ in range(steps):
        t += g.step(t)
        its.append(g.stats().last_parabolic_its)
    out = dict(p=g.get_pressure(), u=g.get_velocity(), c=g.get_concentration(), its=its,
               misses=g.stats().cn_ahead_misses, t=t)
    g.close()
    return o, out


def test_lookahead_leaves_fields_bitwise_unchanged(gpu_lib):
    runs = {}
    for ahead in (-1.0, 0.0, 1.0):
        print("lookahead factor", ahead, flush=True)
        o, runs[ahead] = _run(ahead)
        print("  BiCGStab iterations per step %s, misses %d" % (runs[ahead]["its"], runs[ahead]["misses"]),
              flush=True)
    ref = runs[-1.0]
    assert ref["misses"] == 0
    assert all(k > 1 for k in ref["its"])          # the solves iterate
    assert runs[1.0]["misses"] >= len(ref["its"])  # every solve ended on a queued p-launch
    for ahead in (0.0, 1.0):
        r = runs[ahead]
        assert r["its"] == ref["its"]
        assert r["t"] == ref["t"]
        for f in ("p", "u", "c"):
            assert np.array_equal(r[f], ref[f]), (ahead, f)


def test_lookahead_matches_oracle(gpu_lib):
    """The default factor against the oracle over the same three steps."""
    p = params(N1=48, N2=40, eps=1e-12, para_eps=1e-12, mu_inj=0.1, diffusion=1, alpha_abs=1,
               alpha_l=1e-2, alpha_t=1e-3, d_m=1e-4)
    o, g = pair(p, cn_ahead=1.0)
    q = o.source("one_cell")
    c, pr, v = o.scalar(), o.scalar(), o.velocity()
    g.set_source(q)
    g.set_concentration(c)
    t_ref = t_gpu = 0.0
    for _ in range(3):
        st, dt = o.step(c, pr, v, q, t_ref)
        assert st == 0
        t_ref += dt
        t_gpu += g.step(t_gpu)
    assert abs(t_gpu - t_ref) <= 1e-10 * t_ref
    assert rel_err(o.interior(g.get_concentration()), o.interior(c)) < 1e-9
    assert g.stats().cn_ahead_misses > 0
    g.close()


def _run_pressure(ahead, precision, steps=3):
    p = params(N1=64, N2=48, eps=1e-11, mu_inj=0.1)
    o, g = pair(p, cg_ahead=ahead, precision=precision)
    g.set_source(o.source("one_cell"))
    g.set_concentration(o.scalar())
    t = 0.0
    its = []
    for _ in range(steps):
        t += g.step(t)
        its.append(g.stats().last_its)
    out = dict(p=g.get_pressure(), u=g.get_velocity(), c=g.get_concentration(), its=its,
               misses=g.stats().cg_ahead_misses, t=t)
    g.close()
    return out


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_pressure_lookahead_leaves_fields_bitwise_unchanged(gpu_lib, precision):
    """The pressure solve queues the next V-cycle's first leg before the host
    reads a CG iteration's residual while that residual is far above the
    target; the leg reads only r, so the fields do not depend on the factor
    (1: every iteration ahead, each solve ending with a leg that goes unused)."""
    from impec2d_amd import capi
    prec = capi.PREC_FP64 if precision == "fp64" else capi.PREC_MIXED
    runs = {}
    for ahead in (-1.0, 0.0, 1.0):
        runs[ahead] = _run_pressure(ahead, prec)
        print(precision, "factor", ahead, "CG iterations", runs[ahead]["its"], "misses", runs[ahead]["misses"],
              flush=True)
    ref = runs[-1.0]
    assert ref["misses"] == 0
    assert runs[1.0]["misses"] >= len(ref["its"])
    for ahead in (0.0, 1.0):
        r = runs[ahead]
        assert r["its"] == ref["its"] and r["t"] == ref["t"]
        for f in ("p", "u", "c"):
            assert np.array_equal(r[f], ref[f]), (precision, ahead, f)
