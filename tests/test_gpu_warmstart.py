"""Warm starts: the previous pressure as the MG-CG initial guess
(initial_guess = previous), the previous pressure plus the last step's change
with a minimal-residual step length (initial_guess = extrapolate), and the
current concentration as the dispersion solve's initial guess (cn_guess = 1).  Both keep the stopping rule relative to
||b||, so the solutions agree with the zero-guess (reference) solves to the
solver tolerance, with fewer iterations."""
import numpy as np
import pytest

from helpers import pair, params, rel_err

pytestmark = pytest.mark.gpu


def _run(p, steps, **extra):
    o, g = pair(p, **extra)
    g.set_source(o.source("one_cell"))
    g.set_concentration(o.scalar())
    t, its, pits = 0.0, [], []
    for _ in range(steps):
        t += g.step(t)
        st = g.stats()
        its.append(st.last_its)
        pits.append(st.last_parabolic_its)
    out = (o.interior(g.get_pressure()), o.interior(g.get_concentration()), t, its, pits)
    g.close()
    return out


@pytest.mark.parametrize("n", [64, 256])
def test_previous_pressure_and_current_c_guesses(gpu_lib, n):
    from impec2d_amd import capi
    p = params(N1=n, N2=n, eps=1e-12, para_eps=1e-12, mu_inj=0.1, diffusion=1, alpha_abs=1, alpha_l=1e-3,
               alpha_t=1e-4, d_m=1e-5)
    print("warm start", n, flush=True)
    p0, c0, t0, its0, pits0 = _run(p, 8)
    p1, c1, t1, its1, pits1 = _run(p, 8, initial_guess=capi.GUESS_PREVIOUS, cn_guess=1)
    p2, c2, t2, its2, _ = _run(p, 8, initial_guess=capi.GUESS_EXTRAPOLATE)
    print("  its zero", its0, "previous", its1, "extrapolated", its2, "; parabolic its zero", pits0,
          "current c", pits1, flush=True)
    assert rel_err(p1, p0) < 1e-9 and rel_err(c1, c0) < 1e-9
    assert rel_err(p2, p0) < 1e-9 and rel_err(c2, c0) < 1e-9
    assert abs(t1 - t0) <= 1e-9 * t0 and abs(t2 - t0) <= 1e-9 * t0
    assert sum(its1[1:]) <= sum(its0[1:])
    assert sum(its2[2:]) <= sum(its1[2:]) + 1
    assert sum(pits1) <= sum(pits0) + 1
