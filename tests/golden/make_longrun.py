"""Writes the long-run fixtures of tests/golden/longrun/ from the CPU oracle.

Each fixture runs the oracle's IMPEC step loop from c = 0 to an end time (the
loop of Impec2DSolver::Solve, Impec2DSolver.cpp:339-442, with the last step
clipped to the end time) or for a number of steps, and keeps:
  * the inputs: parameters (JSON), well profile;
  * per step: t, dt, the cumulative produced mass (total_mass_outflux) and
    the pressure solve's iterations;
  * the final c, p and v (with ghosts) of the 64^2 cases.
The GPU tests compare the HIP path's curves (impec2d_amd/recovery.py) with
these, pointwise on the stable cases and within stated tolerances on the
unstable ones.  test_longrun_fixtures.py checks that the oracle still
reproduces the first steps of every fixture bit for bit.  Regenerate with
`python tests/golden/make_longrun.py [case ...]` (a few minutes in all).
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "longrun")
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import pyoracle as po  # noqa: E402

R2 = 1.41421356237
# the quarter-five-spot of the CPU configuration: injector (0,0), producer (1,1),
# one-cell wells of rate 1, 64^2, M = 10, MUSCL with parallel splitting
C1 = dict(N1=64, N2=64, stype=1, mu_res=1.0, mu_inj=0.1, nu=1.0, eps=1e-10, cfl=0.5)
# the shipped production configurations' physics (examples/*.cfg): circle wells,
# M = 41, nu = 2/3; their elliptic/parabolic eps of 1e-14 lies below the FP64
# rounding floor of these systems, so the oracle runs at 1e-12
PAR = dict(N1=141, N2=141, stype=0, x_ub=R2, y_ub=R2, nu=2 / 3.0, mu_res=41.0, mu_inj=1.0, eps=1e-12,
           para_eps=1e-12, cfl=0.5, I=1.0, well_rc=0.05)
DIAG = dict(PAR, N1=100, N2=50, stype=1, x_ub=1.0, y_ub=1.0)
DISP_CFG = dict(diffusion=1, alpha_abs=0, alpha_l=1.0, alpha_t=1.0)

CASES = {
    # the CPU configuration exactly: 200 steps
    "c1_200": (dict(C1), "one_cell", None, 200),
    # stable: M = 1 to breakthrough and beyond
    "m1_t080": (dict(C1, mu_inj=1.0), "one_cell", 0.8, None),
    # unstable: M = 10 to t = 0.8
    "m10_t080": (dict(C1), "one_cell", 0.8, None),
    # stable through strong dispersion: M = 10, alpha_l 0.02, alpha_t 0.005, d_m 1e-3 (absolute)
    "m10_disp_t080": (dict(C1, diffusion=1, alpha_abs=1, alpha_l=0.02, alpha_t=0.005, d_m=1e-3, para_eps=1e-10),
                      "one_cell", 0.8, None),
    # examples/parallel.cfg to its end time 0.4
    "parallel_cfg": (dict(PAR), "circle", 0.4, None),
    # examples/parallel_dispersion.cfg and diagonal_dispersion.cfg to 0.4
    "parallel_dispersion_cfg": (dict(PAR, **DISP_CFG), "circle", 0.4, None),
    "diagonal_dispersion_cfg": (dict(DIAG, **DISP_CFG), "circle", 0.4, None),
}


def run(case, verbose=True):
    kw, profile, tend, steps = CASES[case]
    o = po.Oracle(po.default_params(**kw))
    q = o.source(profile)
    c, p, v = o.scalar(), o.scalar(), o.velocity()
    ts, dts, outs, its = [], [], [], []
    t, k = 0.0, 0
    t0 = time.time()
    while (steps is not None and k < steps) or (tend is not None and t < tend):
        st, dt = o.step(c, p, v, q, t, tend if tend is not None else -1.0)
        if st != 0:
            raise RuntimeError("%s: oracle step %d returned %d" % (case, k, st))
        t += dt
        k += 1
        ts.append(t)
        dts.append(dt)
        outs.append(o.total_mass_outflux)
        its.append(o.last_its)
        if verbose and k % 200 == 0:
            print("  %s: step %d t %.4f (%.0f s)" % (case, k, t, time.time() - t0), flush=True)
    meta = dict(params=kw, profile=profile, tend=tend, steps=steps)
    arr = dict(t=np.array(ts), dt=np.array(dts), outflux=np.array(outs), its=np.array(its))
    if kw["N1"] * kw["N2"] <= 64 * 64:   # final fields of the small cases only (fixture size)
        arr.update(c=c, p=p, v=v)
    return meta, arr


def main(names):
    os.makedirs(OUT, exist_ok=True)
    for case in names or CASES:
        meta, arr = run(case)
        np.savez_compressed(os.path.join(OUT, case + ".npz"), meta=np.array(json.dumps(meta)), **arr)
        print("wrote", case, "steps", arr["t"].size, "t", arr["t"][-1], flush=True)


if __name__ == "__main__":
    main(sys.argv[1:])
