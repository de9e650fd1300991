"""The long-run fixtures (tests/golden/longrun/) and the curve definitions of
impec2d_amd/recovery.py, on the CPU.

* the oracle reproduces the first steps of every fixture bit for bit (the
  whole runs take minutes; make_longrun.py regenerates them);
* the curves follow from their definitions on cases whose answer is known:
  before breakthrough the recovered fraction is the injected pore volume;
  a produced concentration that is constant after a time gives that value
  and a breakthrough at that time; the fixtures' own mass balance holds."""
import json
import os

import numpy as np
import pytest

from impec2d_amd import recovery
from oracle import pyoracle as po

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "longrun")
CASES = sorted(f[:-4] for f in os.listdir(HERE) if f.endswith(".npz")) if os.path.isdir(HERE) else []


def _load(case):
    z = np.load(os.path.join(HERE, case + ".npz"))
    return json.loads(str(z["meta"])), z


def test_fixtures_present():
    assert set(CASES) >= {"c1_200", "m1_t080", "m10_t080", "m10_disp_t080", "parallel_cfg",
                          "parallel_dispersion_cfg", "diagonal_dispersion_cfg"}


@pytest.mark.parametrize("case", CASES)
def test_oracle_reproduces_first_steps(case):
    meta, z = _load(case)
    o = po.Oracle(po.default_params(**meta["params"]))
    q = o.source(meta["profile"])
    c, p, v = o.scalar(), o.scalar(), o.velocity()
    t = 0.0
    tend = meta["tend"] if meta["tend"] is not None else -1.0
    for k in range(3):
        st, dt = o.step(c, p, v, q, t, tend)
        assert st == 0
        t += dt
        assert t == z["t"][k] and dt == z["dt"][k] and o.total_mass_outflux == z["outflux"][k]


@pytest.mark.parametrize("case", CASES)
def test_fixture_runs_reach_their_end(case):
    meta, z = _load(case)
    if meta["tend"] is not None:
        assert z["t"][-1] >= meta["tend"] - 1e-12 and z["t"][-2] < meta["tend"]
    else:
        assert z["t"].size == meta["steps"]
    assert np.all(np.diff(z["t"]) > 0) and np.all(np.diff(z["outflux"]) >= 0)


def _geo(**kw):
    d = dict(N1=10, N2=10, x_lb=0.0, x_ub=1.0, y_lb=0.0, y_ub=1.0, I=1.0, source_type=1, porosity=1.0)
    d.update(kw)
    return d


def test_recovered_fraction_before_breakthrough_is_injected_pore_volume():
    t = np.linspace(0.01, 0.3, 30)
    for st, n_inj in ((1, 1.0), (0, 2.0)):
        geo = _geo(source_type=st, N1=8, N2=5, x_ub=2.0, y_ub=0.5)
        cur = recovery.curves(t, np.zeros_like(t), **geo)
        # injected mass in cell units: t I / (dx dy) per injector; pore volume N1 N2
        dx, dy = 2.0 / 8, 0.5 / 5
        assert np.allclose(cur["injected"], t * n_inj / (dx * dy), rtol=1e-14)
        assert np.allclose(cur["recovered"], t * n_inj / (dx * dy) / 40.0, rtol=1e-14)
        assert np.all(cur["produced_conc"] == 0.0)


def test_constant_produced_concentration_and_breakthrough():
    geo = _geo()
    t = np.linspace(0.01, 1.0, 100)
    rate = 1.0 / (0.1 * 0.1)
    tb, cb = 0.455, 0.3
    inj = t * rate
    out = np.where(t > tb, (t - tb) * rate * cb, 0.0)
    cur = recovery.curves(t, out, **geo)
    after = t > tb + 0.011
    assert np.allclose(cur["produced_conc"][after], cb, rtol=1e-12)
    assert np.allclose(cur["recovered"], (inj - out) / 100.0, rtol=1e-14)
    # the produced concentration rises within the step containing tb: the
    # interpolated crossing of a low threshold lies in that step
    bt = recovery.breakthrough_time(t, cur["produced_conc"], threshold=0.01)
    k = np.searchsorted(t, tb)
    assert t[k - 1] <= bt <= t[k]
    assert recovery.breakthrough_time(t, np.zeros_like(t)) is None


def test_breakthrough_interpolates_linearly():
    t = np.array([0.1, 0.2, 0.3, 0.4])
    cp = np.array([0.0, 0.0, 0.02, 0.5])
    assert abs(recovery.breakthrough_time(t, cp, 0.01) - 0.25) < 1e-15


def test_on_grid_starts_from_zero():
    g = recovery.on_grid([0.5, 1.0], [1.0, 3.0], np.array([0.0, 0.25, 0.75]))
    assert np.allclose(g, [0.0, 0.5, 2.0])


@pytest.mark.parametrize("case", CASES)
def test_fixture_mass_balance(case):
    """The oracle's produced mass never exceeds what was injected, and the
    small cases' final mass closes the balance (Impec2DSolver.cpp:530-540) to
    the accumulated solver tolerance (a dispersion solve to 1e-10 moves mass
    by about that fraction; the runs have thousands of them)."""
    meta, z = _load(case)
    P = po.default_params(**meta["params"])
    geo = dict(N1=P["N1"], N2=P["N2"], x_lb=P["x_lb"], x_ub=P["x_ub"], y_lb=P["y_lb"], y_ub=P["y_ub"], I=P["I"],
               source_type=P["stype"], porosity=P["porosity"])
    cur = recovery.curves(z["t"], z["outflux"], **geo)
    assert np.all(z["outflux"] <= cur["injected"] * (1 + 1e-12))
    if "c" in z:
        o = po.Oracle(P)
        mass = o.mass(z["c"])
        assert abs(mass + z["outflux"][-1] - cur["injected"][-1]) <= 1e-8 * cur["injected"][-1]
