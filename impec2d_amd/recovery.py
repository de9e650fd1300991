"""Recovery and breakthrough curves of an IMPEC run.

Long runs of an unstable displacement (mobility ratio M > 1 with little
dispersion) amplify rounding differences into different finger patterns, so
two correct solvers only agree on integral measures of the run.  These are
the measures, computed from what the step loop already reports after every
step: the time t and the produced mass of injected fluid summed by the source
split, HyperbolicSolver::GetTotalMassOutflux (TRIntegrator.cpp:49-71).

Units are the reference's mass-balance units (Impec2DSolver.cpp:530-540): a
cell's concentration counts 1 per unit of cell volume, so the injected mass
up to t is (t - t0) I / (dx dy), doubled for the `parallel` source type (two
injectors), and the pore volume is N1 N2 porosity.

  injected(t)       (t - t0) I / (dx dy) * n_injectors
  recovered(t)      resident fluid produced / pore volume
                    = (injected(t) - outflux(t)) / (N1 N2 porosity)
                    (the produced volume equals the injected volume: the
                    flow is incompressible and the sources sum to zero)
  produced_conc     per step, the produced injected-fluid mass over the
                    produced volume, d outflux / d injected: the flux-weighted
                    concentration at the sinks
  breakthrough      the first time produced_conc reaches a threshold (0.01
                    by default), interpolated linearly within the step
"""
import numpy as np

SRC_PARALLEL = 0


def injected(t, N1, N2, x_lb, x_ub, y_lb, y_ub, I, source_type, tstart=0.0):
    dx = (x_ub - x_lb) / N1
    dy = (y_ub - y_lb) / N2
    m = (np.asarray(t, dtype=float) - tstart) * I / (dx * dy)
    return m * (2.0 if source_type == SRC_PARALLEL else 1.0)


def curves(t, outflux, N1, N2, x_lb, x_ub, y_lb, y_ub, I, source_type, porosity=1.0, tstart=0.0):
    """Per-step arrays (t after each step, cumulative outflux after each step)
    -> dict with t, injected, recovered, produced_conc (the first entry's
    produced concentration is that of the first step)."""
    t = np.asarray(t, dtype=float)
    out = np.asarray(outflux, dtype=float)
    if t.shape != out.shape or t.ndim != 1 or t.size == 0:
        raise ValueError("t and outflux must be equal-length 1-D series")
    inj = injected(t, N1, N2, x_lb, x_ub, y_lb, y_ub, I, source_type, tstart)
    rec = (inj - out) / (N1 * N2 * porosity)
    d_out = np.diff(np.concatenate([[0.0], out]))
    d_inj = np.diff(np.concatenate([[0.0], inj]))
    with np.errstate(divide="ignore", invalid="ignore"):
        cp = np.where(d_inj > 0, d_out / d_inj, 0.0)
    return {"t": t, "injected": inj, "recovered": rec, "produced_conc": cp}


def breakthrough_time(t, produced_conc, threshold=0.01, tstart=0.0):
    """First time the produced concentration reaches `threshold`, linear in
    the step that crosses it; None when it never does."""
    t = np.asarray(t, dtype=float)
    cp = np.asarray(produced_conc, dtype=float)
    hit = np.nonzero(cp >= threshold)[0]
    if hit.size == 0:
        return None
    k = int(hit[0])
    t0, c0 = (t[k - 1], cp[k - 1]) if k > 0 else (tstart, 0.0)
    if cp[k] == c0:
        return float(t[k])
    return float(t0 + (threshold - c0) / (cp[k] - c0) * (t[k] - t0))


def on_grid(t, values, grid):
    """The step series resampled on a common time grid (piecewise linear;
    the state before the first step is taken as zero)."""
    t = np.concatenate([[0.0], np.asarray(t, dtype=float)])
    v = np.concatenate([[0.0], np.asarray(values, dtype=float)])
    return np.interp(grid, t, v)
