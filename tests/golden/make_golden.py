"""Writes the golden fixtures of tests/golden/ from the CPU oracle.

Each fixture holds the inputs (parameters, source) and the outputs (c, p, v
interiors with ghosts, dt) after 1 and 10 IMPEC steps from c = 0 on a 16^2
grid.  Regenerate with `python tests/golden/make_golden.py`; test_golden.py
checks that the oracle still reproduces them bit for bit and the GPU tests
compare the HIP path against them.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import pyoracle as po  # noqa: E402

CASES = {
    "diag_nu1": dict(N1=16, N2=16, nu=1.0, mu_inj=0.1),
    "diag_nu23": dict(N1=16, N2=16, nu=2 / 3.0, mu_inj=0.1),
    "par_nu1": dict(N1=16, N2=16, nu=1.0, mu_inj=0.1, stype=0, x_ub=2 ** 0.5, y_ub=2 ** 0.5),
    "diag_disp": dict(N1=16, N2=16, nu=1.0, mu_inj=0.1, diffusion=1, alpha_abs=1, alpha_l=1e-3,
                      alpha_t=1e-4, d_m=1e-5),
}
COMMON = dict(eps=1e-13, para_eps=1e-13)


def run(case):
    kw = dict(COMMON, **CASES[case])
    o = po.Oracle(po.default_params(**kw))
    c, p, v, q = o.scalar(), o.scalar(), o.velocity(), o.source("one_cell")
    out = {"q": q}
    t = 0.0
    for k in range(1, 11):
        st, dt = o.step(c, p, v, q, t)
        assert st == 0
        t += dt
        if k in (1, 10):
            out["c%d" % k], out["p%d" % k], out["v%d" % k] = c.copy(), p.copy(), v.copy()
            out["dt%d" % k], out["t%d" % k] = np.array(dt), np.array(t)
    return kw, out


def main():
    for case in CASES:
        kw, out = run(case)
        np.savez_compressed(os.path.join(HERE, case + ".npz"), params=np.array(json.dumps(kw)), **out)
        print("wrote", case)


if __name__ == "__main__":
    main()
