"""Seeded random configurations against the oracle: grid sizes (odd, non-square),
both grids, the ν blend, all five splitting modes, MUSCL and upwind, harmonic
and mean-concentration averaging, porosity, CFL, mobility ratio, one-cell and
circle wells, with and without CN dispersion (relative and absolute
dispersivities, molecular diffusion).  Each case steps 4 times on the GPU and
in the oracle; p, u, v and c agree to TOL relative and every time step to
1e-10 (both solves to 1e-12, so the fields share about 10 digits)."""
import numpy as np
import pytest

from helpers import pair, params, rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-9
STEPS = 4


def _case(seed):
    rng = np.random.default_rng(seed)
    kw = dict(N1=int(rng.integers(8, 81)), N2=int(rng.integers(8, 81)),
              nu=float(rng.choice([1.0, 2 / 3.0, 1 / 3.0, 0.0])),
              stype=int(rng.integers(0, 2)), dim_split=int(rng.integers(0, 5)), hyper=int(rng.integers(0, 2)),
              avgproc=int(rng.integers(0, 2)), porosity=float(rng.uniform(0.5, 1.0)),
              cfl=float(rng.uniform(0.2, 0.9)), mu_inj=float(1.0 / rng.uniform(1.0, 50.0)),
              x_ub=float(rng.uniform(0.5, 2.0)), y_ub=float(rng.uniform(0.5, 2.0)),
              eps=1e-12, para_eps=1e-12)
    if rng.random() < 0.5:
        kw.update(diffusion=1, alpha_abs=int(rng.integers(0, 2)))
        if kw["alpha_abs"]:
            kw.update(alpha_l=float(rng.uniform(1e-3, 2e-2)), alpha_t=float(rng.uniform(1e-4, 5e-3)),
                      d_m=float(rng.uniform(0.0, 1e-3)))
        else:
            kw.update(alpha_l=float(rng.uniform(0.1, 1.0)), alpha_t=float(rng.uniform(0.01, 0.5)))
    profile = "circle" if rng.random() < 0.3 else "one_cell"
    return kw, profile


SEEDS = list(range(4100, 4116))


@pytest.mark.parametrize("seed", SEEDS)
def test_random_configuration_steps(gpu_lib, seed):
    kw, profile = _case(seed)
    print("random case", seed, profile, kw, flush=True)
    p = params(**kw)
    o, g = pair(p)
    q = o.source(profile)
    c, pr, v = o.scalar(), o.scalar(), o.velocity()
    g.set_source(q)
    g.set_concentration(c)
    t_ref = t_gpu = 0.0
    for k in range(STEPS):
        st, dt_ref = o.step(c, pr, v, q, t_ref)
        assert st == 0
        t_ref += dt_ref
        dt_gpu = g.step(t_gpu)
        t_gpu += dt_gpu
        assert abs(dt_gpu - dt_ref) <= 1e-10 * dt_ref, (k, dt_gpu, dt_ref)
    ep = rel_err(o.interior(g.get_pressure()), o.interior(pr))
    ev = rel_err(g.get_velocity(), v)
    ec = rel_err(o.interior(g.get_concentration()), o.interior(c))
    print("  after %d steps: p %.2e v %.2e c %.2e" % (STEPS, ep, ev, ec), flush=True)
    assert ep < TOL and ev < TOL and ec < TOL
    g.close()
