"""Long runs on the GPU against the oracle's long-run fixtures
(tests/golden/longrun/, written by make_longrun.py).

Stable cases (M = 1, strong dispersion) and the CPU configuration's 200 steps
are compared pointwise at the end: p, u, v within TOL_PV relative, c within
TOL_C relative (the two solvers' pressures differ by their solve tolerance,
1e-10, at every step, and the MUSCL limiter's branches turn such differences
into concentration differences of about 1e-9 over thousands of steps), every
step's t within 1e-10 relative.  Unstable cases (M = 10 without dispersion,
and the shipped production configurations with M = 41) amplify rounding
differences into different fingers, so they are judged by the curves of
impec2d_amd/recovery.py:
  * recovered fraction of the resident fluid on a common time grid:
    max |difference| <= TOL_RECOVERED (fraction of the pore volume);
  * breakthrough time (produced concentration reaching 0.01):
    relative difference <= TOL_BREAKTHROUGH.
The shipped configurations run unchanged through impec2d_run --config to
their end time (its --recovery output carries the curves)."""
import json
import os
import subprocess

import numpy as np
import pytest

from impec2d_amd import recovery
from oracle import pyoracle as po
from tests.helpers import gpu_params, rel_err

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIX = os.path.join(ROOT, "tests", "golden", "longrun")
RUN = os.path.join(ROOT, "impec2d_amd", "impec2d_run")
CFG = os.path.join(ROOT, "tests", "golden", "configs")

TOL_PV = 1e-10
TOL_C = 1e-8
TOL_RECOVERED = 2e-3
TOL_BREAKTHROUGH = 0.02
STABLE = ["c1_200", "m1_t080", "m10_disp_t080"]
UNSTABLE = ["m10_t080"]
CONFIGS = {"parallel_cfg": "parallel.cfg", "parallel_dispersion_cfg": "parallel_dispersion.cfg",
           "diagonal_dispersion_cfg": "diagonal_dispersion.cfg"}


def _say(*a):
    print(*a, flush=True)


def _load(case):
    z = np.load(os.path.join(FIX, case + ".npz"))
    return json.loads(str(z["meta"])), z


def _geometry(kw):
    P = po.default_params(**kw)
    return dict(N1=P["N1"], N2=P["N2"], x_lb=P["x_lb"], x_ub=P["x_ub"], y_lb=P["y_lb"], y_ub=P["y_ub"], I=P["I"],
                source_type=P["stype"], porosity=P["porosity"])


def _gpu_run(meta):
    """The step loop of Impec2DSolver::Solve on the GPU through the C ABI, with
    the same end-time clipping as the oracle's loop."""
    from impec2d_amd import capi
    kw = meta["params"]
    P = po.default_params(**kw)
    o = po.Oracle(P)
    g = capi.Context(gpu_params(P))
    g.set_source(o.source(meta["profile"]))
    g.set_concentration(o.scalar())
    tend, steps = meta["tend"], meta["steps"]
    ts, outs = [], []
    t, k = 0.0, 0
    while (steps is not None and k < steps) or (tend is not None and t < tend):
        t += g.step(t, tend if tend is not None else -1.0)
        k += 1
        ts.append(t)
        outs.append(g.stats().total_mass_outflux)
    return g, o, np.array(ts), np.array(outs)


def _compare_curves(geo, t_ref, out_ref, t_gpu, out_gpu, tend):
    ref = recovery.curves(t_ref, out_ref, **geo)
    gpu = recovery.curves(t_gpu, out_gpu, **geo)
    grid = np.linspace(0.0, min(t_ref[-1], t_gpu[-1]), 401)
    d_rec = np.abs(recovery.on_grid(t_gpu, gpu["recovered"], grid) - recovery.on_grid(t_ref, ref["recovered"], grid))
    bt_ref = recovery.breakthrough_time(t_ref, ref["produced_conc"])
    bt_gpu = recovery.breakthrough_time(t_gpu, gpu["produced_conc"])
    return d_rec.max(), bt_ref, bt_gpu, ref, gpu


@pytest.mark.parametrize("case", STABLE)
def test_stable_long_run_pointwise(gpu_lib, case):
    meta, z = _load(case)
    _say("long run (stable)", case, "steps", z["t"].size)
    g, o, ts, outs = _gpu_run(meta)
    assert ts.size == z["t"].size
    assert np.abs(ts - z["t"]).max() <= 1e-10 * z["t"][-1]
    ec = rel_err(o.interior(g.get_concentration()), o.interior(z["c"]))
    ep = rel_err(o.interior(g.get_pressure()), o.interior(z["p"]))
    ev = rel_err(g.get_velocity(), z["v"])
    _say("  end t %.4f: c %.2e p %.2e v %.2e" % (ts[-1], ec, ep, ev))
    assert ec < TOL_C and ep < TOL_PV and ev < TOL_PV
    scale = max(np.abs(z["outflux"]).max(), 1e-300)
    assert np.abs(outs - z["outflux"]).max() <= 1e-9 * scale
    g.close()


@pytest.mark.parametrize("case", UNSTABLE)
def test_unstable_long_run_curves(gpu_lib, case):
    meta, z = _load(case)
    _say("long run (unstable)", case, "steps", z["t"].size)
    g, o, ts, outs = _gpu_run(meta)
    geo = _geometry(meta["params"])
    d_rec, bt_ref, bt_gpu, ref, _ = _compare_curves(geo, z["t"], z["outflux"], ts, outs, meta["tend"])
    ec = rel_err(o.interior(g.get_concentration()), o.interior(z["c"]))
    _say("  steps gpu %d oracle %d; recovered max diff %.2e; breakthrough gpu %s oracle %s; pointwise c %.2e"
         % (ts.size, z["t"].size, d_rec, bt_gpu, bt_ref, ec))
    assert bt_ref is not None and bt_gpu is not None
    assert d_rec <= TOL_RECOVERED
    assert abs(bt_gpu - bt_ref) <= TOL_BREAKTHROUGH * bt_ref
    assert ref["recovered"][-1] > 0.3   # the run goes well past breakthrough
    g.close()


@pytest.mark.parametrize("case", sorted(CONFIGS))
def test_shipped_configuration_to_end_time(gpu_lib, case, tmp_path):
    """The configuration file unchanged to its end time 0.4 (its eps of 1e-14
    lies below the FP64 rounding floor of these systems: the solves stop at
    that floor, which they report)."""
    meta, z = _load(case)
    out = tmp_path / "out"
    log = tmp_path / "log"
    out.mkdir()
    log.mkdir()
    rec = tmp_path / "recovery.txt"
    cmd = [RUN, "--config", os.path.join(CFG, CONFIGS[case]), "--quiet", "--results-dir", str(out), "--log-dir",
           str(log), "--recovery", str(rec)]
    _say("run", CONFIGS[case], "to its end time")
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    a = np.loadtxt(str(rec), ndmin=2)
    t_gpu, out_gpu = a[:, 1], a[:, 3]
    assert abs(t_gpu[-1] - 0.4) <= 1e-12
    geo = _geometry(meta["params"])
    d_rec, bt_ref, bt_gpu, ref, gpu = _compare_curves(geo, z["t"], z["outflux"], t_gpu, out_gpu, 0.4)
    # the driver's own recovered fraction equals the module's
    assert np.abs(a[:, 5] - gpu["recovered"]).max() <= 1e-12
    _say("  steps gpu %d oracle %d; recovered at 0.4: gpu %.5f oracle %.5f (max diff %.2e); breakthrough gpu %s "
         "oracle %s" % (t_gpu.size, z["t"].size, gpu["recovered"][-1], ref["recovered"][-1], d_rec, bt_gpu, bt_ref))
    assert d_rec <= TOL_RECOVERED
    if bt_ref is not None or bt_gpu is not None:
        assert bt_ref is not None and bt_gpu is not None
        assert abs(bt_gpu - bt_ref) <= TOL_BREAKTHROUGH * bt_ref
    diag = np.loadtxt(str(out / "diagnostics.dat"), ndmin=2)
    assert diag.shape[0] == t_gpu.size
    assert np.abs(diag[:, 5]).max() < 1e-8   # mass balance at every step
