"""The shipped configurations run unchanged through the driver
(impec2d_run --config): circle wells, nu = 2/3, Crank-Nicolson dispersion,
both grids (the diagonal one with dx != dy).  After three steps the fields
agree with the oracle stepped with the same parameters; the output files of
Impec2DSolver::LogResultsMatch / WriteDiagnostics are written in the
reference's formats, and the projection of the parallel grid onto the
diagonal grid matches an independent bilinear interpolation of the dumped
field."""
import math
import os
import subprocess

import numpy as np
import pytest

from helpers import rel_err
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUN = os.path.join(ROOT, "impec2d_amd", "impec2d_run")
CFG = os.path.join(ROOT, "tests", "golden", "configs")
R2 = 1.41421356237


def _say(*a):
    print(*a, flush=True)


def _run(name, tmp_path, steps):
    out = tmp_path / "out"
    log = tmp_path / "log"
    out.mkdir()
    log.mkdir()
    dump = tmp_path / "fields.bin"
    cmd = [RUN, "--config", os.path.join(CFG, name), "--steps", str(steps), "--quiet", "--results-dir", str(out),
           "--log-dir", str(log), "--dump", str(dump)]
    _say("run", " ".join(cmd[1:3]))
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return out, log, np.fromfile(str(dump), dtype=np.float64)


def _oracle(N1, N2, stype, ext, steps):
    p = po.default_params(N1=N1, N2=N2, stype=stype, x_ub=ext, y_ub=ext, nu=2 / 3.0, mu_res=41.0, mu_inj=1.0,
                          diffusion=1, alpha_abs=0, alpha_l=1.0, alpha_t=1.0, eps=1e-14, para_eps=1e-14, cfl=0.5,
                          I=1.0, well_rc=0.05, tfinal=0.4)
    o = po.Oracle(p)
    q = o.source("circle")
    c, pr, v = o.scalar(), o.scalar(), o.velocity()
    t = 0.0
    for _ in range(steps):
        st, dt = o.step(c, pr, v, q, t)
        assert st == 0
        t += dt
    return o, c, pr, v, t


def _split(raw, N1, N2):
    ns, nv = (N1 + 4) * (N2 + 4), (N1 + 5) * (N2 + 5) * 2
    c = raw[:ns].reshape(N1 + 4, N2 + 4)
    p = raw[ns:2 * ns].reshape(N1 + 4, N2 + 4)
    v = raw[2 * ns:2 * ns + nv].reshape(N1 + 5, N2 + 5, 2)
    return c, p, v


def _plot_values(path):
    rows = [ln.split() for ln in open(path) if ln.strip()]
    return np.array([[float(x) for x in r] for r in rows])


def test_parallel_dispersion_configuration(gpu_lib, tmp_path):
    N, steps = 141, 3
    out, log, raw = _run("parallel_dispersion.cfg", tmp_path, steps)
    c, p, v = _split(raw, N, N)
    o, oc, op, ov, t = _oracle(N, N, 0, R2, steps)
    ec, ep, ev = (rel_err(o.interior(c), o.interior(oc)), rel_err(o.interior(p), o.interior(op)), rel_err(v, ov))
    _say("  vs oracle after %d steps: c %.2e p %.2e v %.2e" % (steps, ec, ep, ev))
    assert ec < 1e-9 and ep < 1e-9 and ev < 1e-9
    # diagnostics.dat: counter, ITS_e, t, dt, cells/dt and the three checks
    diag = np.loadtxt(str(out / "diagnostics.dat"))
    assert diag.shape == (steps, 8) and list(diag[:, 0]) == [1, 2, 3]
    assert abs(diag[-1, 2] - t) <= 1e-5 * t   # 6 significant digits, as the reference writes them
    # logged at step 1 (log_skip_steps = 30 logs step 1), final fields projected
    dN = int(round(N / math.sqrt(2)))
    for f in (log / "concentration_1.dat", out / "concentration.dat", out / "pressure.dat"):
        vals = _plot_values(str(f))
        assert vals.shape == (dN * dN, 3), f
    # the projection against a bilinear interpolation of the dumped c
    vals = _plot_values(str(out / "concentration.dat"))
    h = R2 / N
    d_h = R2 / (math.sqrt(2.0) * dN)
    a = -math.pi / 4
    ref = np.empty(dN * dN)
    for k in range(dN * dN):
        i, j = divmod(k, dN)
        xd, yd = (i + 0.5) * d_h, (j + 0.5) * d_h
        x = math.cos(a) * xd - math.sin(a) * yd
        y = abs(math.sin(a) * xd + math.cos(a) * yd)
        # ParallelProjector::GetFromParallel pairs cell floor(x / dx) with the next
        # one (ParallelProjector.cpp:100-141), extrapolating on the lower half
        i0, j0 = int(math.floor(x / h)), int(math.floor(y / h))
        fx, fy = x / h - 0.5 - i0, y / h - 0.5 - j0
        cc = lambda ii, jj: c[ii + 2, jj + 2]
        ref[k] = ((1 - fx) * (1 - fy) * cc(i0, j0) + fx * (1 - fy) * cc(i0 + 1, j0) + (1 - fx) * fy * cc(i0, j0 + 1)
                  + fx * fy * cc(i0 + 1, j0 + 1))
    assert np.allclose(vals[:, 0], (np.arange(dN * dN) // dN + 0.5) * d_h, rtol=1e-5)
    assert np.allclose(vals[:, 2], ref, rtol=2e-5, atol=1e-6 * max(1e-300, np.abs(ref).max()))


def test_diagonal_dispersion_configuration(gpu_lib, tmp_path):
    N1, N2, steps = 100, 50, 3
    out, log, raw = _run("diagonal_dispersion.cfg", tmp_path, steps)
    c, p, v = _split(raw, N1, N2)
    o, oc, op, ov, t = _oracle(N1, N2, 1, 1.0, steps)
    ec, ep, ev = (rel_err(o.interior(c), o.interior(oc)), rel_err(o.interior(p), o.interior(op)), rel_err(v, ov))
    _say("  vs oracle after %d steps: c %.2e p %.2e v %.2e" % (steps, ec, ep, ev))
    assert ec < 1e-9 and ep < 1e-9 and ev < 1e-9
    cv = _plot_values(str(out / "concentration.dat"))
    assert cv.shape == (N1 * N2, 3)
    assert np.allclose(cv[:, 2], o.interior(c).ravel(), rtol=1e-5, atol=1e-12)
    assert _plot_values(str(out / "velocity_u.dat")).shape == ((N1 + 1) * N2, 3)
    assert _plot_values(str(out / "velocity_v.dat")).shape == (N1 * (N2 + 1), 3)
    assert np.loadtxt(str(out / "diagnostics.dat")).shape == (steps, 8)
