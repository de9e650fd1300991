#!/usr/bin/env python3
"""IMPEC cell-updates/s on one MI355X (gfx950).

Workload (the plain FP64 GPU configuration at 4096^2): quarter-five-spot on the
unit square (reference source_type "diagonal": injector at cell (0,0),
producer at (N-1,N-1), one-cell wells of rate 1), K = 1, porosity 1,
mobility ratio M = 10 with quarter-power mixing (mu_res = 1, mu_inj = 0.1),
harmonic face mobilities, the 5-point pressure operator (nu = 1), MUSCL
transport with "parallel" dimensional splitting, CFL 0.5, Peaceman
dispersion (d_m = 1e-5, alpha_l = 1e-3, alpha_t = 1e-4, absolute) by
Crank-Nicolson, pressure and dispersion solves to relative residual 1e-10,
FP64 throughout.  Synthetic inputs: c = 0 at t = 0.

A step is one pass of Impec2DSolver::Solve's loop body (impec_step): ghost
fill, pressure solve, Darcy velocities, CFL time step, source / dispersion /
advection / source / dispersion splitting.  value = N^2 * steps / seconds with
all state resident in HBM.

Prints one JSON line.  --gpus N with N > 1 exits: the path runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def workload_params(capi, n, precision="fp64", guess="zero", mg_rows=0, cn_guess=0, grid="diagonal"):
    """The headline physics at n^2.  grid "diagonal": the unit square with the
    injector at (0,0) and the producer at (1,1) (reference source_type
    "diagonal"); "rotated": the 45-degree rotated square of side sqrt(2) with
    injectors at two opposite corners and producers at the other two
    (reference source_type "parallel")."""
    ext = 1.0 if grid == "diagonal" else 2 ** 0.5
    return capi.default_params(
        N1=n, N2=n, source_type=capi.SRC_DIAGONAL if grid == "diagonal" else capi.SRC_PARALLEL,
        average_procedure=capi.AVG_HARMONIC,
        dim_splitting=capi.SPLIT_PARALLEL, hyper_method=capi.HYPER_MUSCL, diffusion=1,
        para_method=capi.PARA_CN, alpha_absolute=1, x_lb=0.0, x_ub=ext, y_lb=0.0, y_ub=ext,
        nu=1.0, mu_res=1.0, mu_inj=0.1, porosity=1.0, cfl=0.5, inj_fluid_conc=1.0,
        alpha_l=1e-3, alpha_t=1e-4, d_m=1e-5, eps=1e-10, para_eps=1e-10, I=1.0,
        precision={"mixed": capi.PREC_MIXED, "mixed32": capi.PREC_MIXED32}.get(precision, capi.PREC_FP64),
        initial_guess=capi.GUESS_PREVIOUS if guess == "previous" else capi.GUESS_ZERO, mg_rows=mg_rows,
        cn_guess=cn_guess)


def one_cell_source(n):
    import numpy as np
    q = np.zeros(n * n)
    val = 1.0 / ((1.0 / n) * (1.0 / n))   # I / (dx dy), Impec2DSolver.cpp:246
    q[0] = val
    q[n * n - 1] = -val
    return q


def lognormal_params(capi, n, precision, guess):
    """The mixed-precision workload: ln K Gaussian (sigma 1, squared-exponential
    correlation length 0.05, seed 1), M = 20, no dispersion, otherwise as the
    headline workload."""
    p = workload_params(capi, n, precision, guess)
    p.diffusion = 0
    p.mu_inj = 1.0 / 20
    return p


def train_guess_network(capi, n=512, steps=120, crops=32, iters=3000):
    """The learned initial guess, trained on pairs from a lognormal-K run at n^2
    (the network is resolution-agnostic: features and target are normalised
    per field and the increment's step length is fitted in every solve)."""
    from impec2d_amd import fields
    net = capi.Network(layers=8, capacity=8192, crop=64, seed=3)
    ctx = capi.Context(lognormal_params(capi, n, "fp64", "previous"))
    ctx.set_permeability(fields.lognormal_permeability(n, n, sigma=1.0, corr_len=0.05, seed=11), None)
    ctx.set_source(one_cell_source(n))
    ctx.set_warmstart(net)
    t = 0.0
    for k in range(steps):
        t += ctx.step(t)
        if k >= 1:
            net.collect(ctx, crops)
    ctx.close()
    losses = [net.train(iters // 2, batch=32, lr=1e-3), net.train(iters - iters // 2, batch=32, lr=3e-4)]
    return net, {"train_size": n, "pairs": net.pairs(), "iterations": iters, "final_loss": losses[-1]}


def modes(capi, n, warmup=3, steps=8):
    """Plain FP64 against mixed precision and the warm starts on the lognormal-K
    workload at n^2: ms/step and V-cycles/step of each."""
    import numpy as np
    from impec2d_amd import fields
    K = fields.lognormal_permeability(n, n, sigma=1.0, corr_len=0.05, seed=1)
    net, train_info = train_guess_network(capi)
    ref_fields = None
    res = {"workload": "lognormal K (sigma 1, corr. length 0.05), M=20, no dispersion, %d^2, rel-res 1e-10" % n,
           "learned_training": train_info}
    for name, prec, guess in [("fp64_zero", "fp64", "zero"), ("fp64_previous", "fp64", "previous"),
                              ("mixed_zero", "mixed", "zero"),
                              ("mixed_previous", "mixed", "previous"), ("mixed_extrapolated", "mixed", "extrapolate"),
                              ("mixed_learned", "mixed", "learned"), ("mixed32_zero", "mixed32", "zero"),
                              ("mixed32_previous", "mixed32", "previous"), ("mixed32_learned", "mixed32", "learned"),
                              ("mixed_previous_transport32", "mixed", "previous")]:
        p = lognormal_params(capi, n, prec, "zero")
        p.transport32 = 1 if name.endswith("transport32") else 0
        p.initial_guess = {"zero": capi.GUESS_ZERO, "previous": capi.GUESS_PREVIOUS,
                           "extrapolate": capi.GUESS_EXTRAPOLATE, "learned": capi.GUESS_LEARNED}[guess]
        ctx = capi.Context(p)
        ctx.set_permeability(K, None)
        ctx.set_source(one_cell_source(n))
        if guess == "learned":
            ctx.set_warmstart(net)
        t = 0.0
        for _ in range(warmup):
            t += ctx.step(t)
        ctx.synchronize()
        s0 = ctx.stats()
        t0 = time.perf_counter()
        for _ in range(steps):
            t += ctx.step(t)
        ctx.synchronize()
        el = time.perf_counter() - t0
        s1 = ctx.stats()
        res[name] = {"ms_per_step": 1e3 * el / steps, "cell_updates_per_s": n * n * steps / el,
                     "vcycles_per_step": (s1.total_its - s0.total_its) / steps,
                     "last_relres": s1.last_relres}
        # the error per field against the plain FP64 run after the same steps
        fields_now = (ctx.get_pressure()[2:-2, 2:-2], ctx.get_velocity(), ctx.get_concentration()[2:-2, 2:-2])
        if name == "fp64_zero":
            ref_fields = fields_now
        elif name in ("mixed_previous", "mixed32_previous", "mixed_previous_transport32"):
            err = {}
            for key, a, b in (("p", fields_now[0], ref_fields[0]), ("u", fields_now[1][..., 0], ref_fields[1][..., 0]),
                              ("v", fields_now[1][..., 1], ref_fields[1][..., 1]),
                              ("c", fields_now[2], ref_fields[2])):
                err[key] = float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
            res[name]["error_vs_fp64"] = err
        del fields_now
        if guess in ("learned", "extrapolate"):
            res[name]["last_step_lengths"] = {"network_increment": s1.last_guess_alpha,
                                              "last_change": s1.last_guess_gamma}
        ctx.close()
    net.close()
    res["speedup_mixed_learned_vs_fp64"] = res["fp64_zero"]["ms_per_step"] / res["mixed_learned"]["ms_per_step"]
    res["speedup_mixed_extrapolated_vs_fp64"] = (res["fp64_zero"]["ms_per_step"]
                                                 / res["mixed_extrapolated"]["ms_per_step"])
    res["speedup_mixed_previous_vs_fp64"] = res["fp64_zero"]["ms_per_step"] / res["mixed_previous"]["ms_per_step"]
    res["speedup_mixed_vs_fp64"] = res["fp64_zero"]["ms_per_step"] / res["mixed_zero"]["ms_per_step"]
    res["speedup_mixed_previous_transport32_vs_fp64"] = (res["fp64_zero"]["ms_per_step"]
                                                         / res["mixed_previous_transport32"]["ms_per_step"])
    # the warm start's gain alone (FP64) and the precision's gain alone (previous guess)
    res["speedup_fp64_previous_vs_fp64"] = res["fp64_zero"]["ms_per_step"] / res["fp64_previous"]["ms_per_step"]
    res["speedup_mixed_previous_vs_fp64_previous"] = (res["fp64_previous"]["ms_per_step"]
                                                      / res["mixed_previous"]["ms_per_step"])
    for g in ("zero", "previous", "learned"):
        res["speedup_mixed32_%s_vs_fp64" % g] = (res["fp64_zero"]["ms_per_step"]
                                                 / res["mixed32_%s" % g]["ms_per_step"])
    return res


def rotated(capi, n, warmup=3, steps=10):
    """The headline physics on the 45-degree rotated grid (reference
    source_type "parallel": the rotated square of side sqrt(2), injectors at
    two opposite corners, producers at the other two) at n^2, FP64."""
    ctx = capi.Context(workload_params(capi, n, grid="rotated"))
    ctx.set_source(corner_wells_source(n))
    t = 0.0
    for _ in range(warmup):
        t += ctx.step(t)
    ctx.synchronize()
    s0 = ctx.stats()
    t0 = time.perf_counter()
    for _ in range(steps):
        t += ctx.step(t)
    ctx.synchronize()
    el = time.perf_counter() - t0
    s1 = ctx.stats()
    ctx.close()
    return {"workload": "rotated square %d^2, two injectors and two producers at the corners, otherwise the "
                        "headline workload (M=10, Peaceman CN dispersion, FP64, zero guesses)" % n,
            "ms_per_step": 1e3 * el / steps, "cell_updates_per_s": n * n * steps / el,
            "vcycles_per_step": (s1.total_its - s0.total_its) / steps, "parabolic_its_last": s1.last_parabolic_its,
            "pressure_relres_last": s1.last_relres, "parabolic_relres_last": s1.last_parabolic_relres}


def fingered_front(n, radius=0.4, amp=0.05, width=0.03, modes=7):
    """A developed displacement front on the unit square: c = 1 behind a
    quarter circle of the given radius around the injector at (0, 0), with
    `modes` fingers of amplitude `amp`, smeared over `width` (ghosted array)."""
    import numpy as np
    x = (np.arange(n) + 0.5) / n
    X, Y = np.meshgrid(x, x, indexing="ij")
    r, th = np.hypot(X, Y), np.arctan2(Y, X + 1e-300)
    c = 0.5 * (1.0 - np.tanh((r - radius - amp * np.sin(modes * th)) / width))
    out = np.zeros((n + 4, n + 4))
    out[2:-2, 2:-2] = c
    return out


def late_time(capi, n, steps=10, warmup=3):
    """The headline workload at a developed front.  At n^2 = 4096^2 its first
    2000 steps reach t = 1.5e-4 and sweep 0.01% of the cells, so the late
    state is set directly: a fingered front at radius 0.4 (about 12% of the
    domain swept), then the same timing as the headline line, with the solver
    iteration counts against those of the first steps."""
    import numpy as np
    ctx = capi.Context(workload_params(capi, n))
    ctx.set_source(one_cell_source(n))
    ctx.set_concentration(fingered_front(n))
    t = 0.0
    for _ in range(warmup):
        t += ctx.step(t)
    ctx.synchronize()
    s0 = ctx.stats()
    pi = []
    t0 = time.perf_counter()
    for _ in range(steps):
        t += ctx.step(t)
        pi.append(ctx.stats().last_parabolic_its)
    ctx.synchronize()
    el = time.perf_counter() - t0
    s1 = ctx.stats()
    swept = float(np.mean(ctx.get_concentration()[2:-2, 2:-2] > 0.5))
    ctx.close()
    return {"workload": "the headline workload (quarter-five-spot %d^2, M=10, Peaceman CN dispersion, FP64, zero "
                        "guesses) from a fingered front at radius 0.4 (7 fingers, amplitude 0.05, width 0.03)" % n,
            "swept_fraction": swept, "ms_per_step": 1e3 * el / steps, "cell_updates_per_s": n * n * steps / el,
            "vcycles_per_step": (s1.total_its - s0.total_its) / steps,
            "parabolic_its_per_solve": float(np.mean(pi)),
            "pressure_relres_last": s1.last_relres, "parabolic_relres_last": s1.last_parabolic_relres}


def corner_wells_source(n, I=1.0, extent=2 ** 0.5):
    """Reference source_type "parallel" on the rotated square: injectors at
    cells (0,0) and (N-1,N-1), producers at (0,N-1) and (N-1,0), rate I each
    (Util.cpp:565-573)."""
    import numpy as np
    h = extent / n
    val = I / (h * h)
    q = np.zeros(n * n)
    q[0] = val
    q[n * n - 1] = val
    q[n - 1] = -val
    q[(n - 1) * n] = -val
    return q


def ensemble(capi, members=256, warmup=2, steps=10, loop_steps=1, seed=2024, graphs=True):
    """Batched ensemble: `members` realisations with sizes drawn uniformly from
    {128, 256, 512, 1024}^2, each with its own lognormal K (sigma uniform in
    [0.5, 2], correlation length 0.05, own seed) and mobility ratio M uniform in
    [2, 50], the two grids mixed, no dispersion, FP32 V-cycles in the FP64 CG,
    previous-pressure guess, rel-res 1e-10.  The finest-level chunk heights are
    autotuned first.  Reports the ensemble's step time with batched launches
    and graph replays, and the time of the same members stepped one after
    another on their own contexts."""
    import numpy as np
    from impec2d_amd import fields
    rng = np.random.default_rng(seed)
    sizes = rng.choice([128, 256, 512, 1024], members)
    sigmas = rng.uniform(0.5, 2.0, members)
    Ms = rng.uniform(2.0, 50.0, members)
    grids = rng.integers(0, 2, members)   # 0: unit square, injector/producer; 1: rotated, corner wells
    plist = []
    for k in range(members):
        n, ext = int(sizes[k]), (1.0 if grids[k] == 0 else 2 ** 0.5)
        plist.append(capi.default_params(
            N1=n, N2=n, source_type=capi.SRC_DIAGONAL if grids[k] == 0 else capi.SRC_PARALLEL,
            average_procedure=capi.AVG_HARMONIC, dim_splitting=capi.SPLIT_PARALLEL, hyper_method=capi.HYPER_MUSCL,
            diffusion=0, x_lb=0.0, x_ub=ext, y_lb=0.0, y_ub=ext, nu=1.0, mu_res=1.0, mu_inj=1.0 / Ms[k],
            porosity=1.0, cfl=0.5, inj_fluid_conc=1.0, eps=1e-10, I=1.0, precision=capi.PREC_MIXED,
            initial_guess=capi.GUESS_PREVIOUS))
    t0 = time.perf_counter()
    ens = capi.Ensemble(plist)
    for k in range(members):
        n, ext = int(sizes[k]), (1.0 if grids[k] == 0 else 2 ** 0.5)
        m = ens.members[k]
        m.set_permeability(fields.lognormal_permeability(n, n, sigma=sigmas[k], corr_len=0.05, seed=1000 + k,
                                                         x_extent=ext, y_extent=ext), None)
        m.set_source(one_cell_source(n) if grids[k] == 0 else corner_wells_source(n))
    setup_s = time.perf_counter() - t0
    if not graphs:
        ens.set_graphs(False)
    rows = ens.autotune(reps=2)
    cells = int(np.sum(sizes.astype(np.int64) ** 2))
    for _ in range(warmup):
        ens.step()
    its_total = 0
    t0 = time.perf_counter()
    for _ in range(steps):
        _, its = ens.step()
        its_total += int(its.sum())
    el = time.perf_counter() - t0
    # the same members, one context after another (no batching, no graphs)
    tl = [0.0] * members
    t1 = time.perf_counter()
    for _ in range(loop_steps):
        for k, m in enumerate(ens.members):
            tl[k] += m.step(tl[k])
    ens.members[0].synchronize()
    el_loop = time.perf_counter() - t1
    ens.close()
    tuned = {}
    for k in range(members):
        tuned[str(int(sizes[k]))] = int(rows[k])
    return {
        "workload": "%d realisations, sizes uniform from {128,256,512,1024}^2, lognormal K (sigma U[0.5,2], "
                    "corr. length 0.05), M U[2,50], both grids, no dispersion, mixed precision, previous-p "
                    "guess, rel-res 1e-10; %d timed steps after %d warm-up" % (members, steps, warmup),
        "members": members, "cells": cells,
        "members_by_size": {str(s): int(np.sum(sizes == s)) for s in (128, 256, 512, 1024)},
        "ms_per_step": 1e3 * el / steps, "cell_updates_per_s": cells * steps / el,
        "cg_iterations_per_member_step": its_total / (steps * members),
        "per_member_loop_ms_per_step": 1e3 * el_loop / loop_steps,
        "speedup_vs_per_member_loop": (el_loop / loop_steps) / (el / steps),
        "autotuned_rows_by_size": tuned, "setup_s": setup_s, "graphs": bool(graphs),
    }


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _oracle_rate(po, P, steps):
    """cell-updates/s of the oracle's IMPEC step on P after one untimed step."""
    o = po.Oracle(P)
    c, p, v, q = o.scalar(), o.scalar(), o.velocity(), o.source()
    t = 0.0
    st, dt = o.step(c, p, v, q, t)
    t += dt
    t0 = time.perf_counter()
    for _ in range(steps):
        st, dt = o.step(c, p, v, q, t)
        t += dt
    el = time.perf_counter() - t0
    return P["N1"] * P["N2"] * steps / el, el


def cpu_baseline(n=1024, steps=4, c1_steps=100):
    """Times the C oracle (FP64 restatement of the reference algorithm, its
    per-cell loops under OpenMP) on the host cores: `steps` steps of the
    headline workload at n^2 and `c1_steps` steps of the CPU configuration
    (64^2, M=10, no dispersion), each with all cores and with one."""
    import os
    from oracle import pyoracle as po
    ncores = len(os.sched_getaffinity(0))
    # the process's CPU share: OMP_NUM_THREADS when the host sets it
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and 0 < int(env) < ncores:
        ncores = int(env)
    head = po.default_params(N1=n, N2=n, mu_res=1.0, mu_inj=0.1, nu=1.0, eps=1e-10, para_eps=1e-10,
                             diffusion=1, alpha_abs=1, alpha_l=1e-3, alpha_t=1e-4, d_m=1e-5)
    c1 = po.default_params(N1=64, N2=64, mu_res=1.0, mu_inj=0.1, nu=1.0, eps=1e-10)
    res = {}
    for threads in (ncores, 1):
        used = po.set_threads(threads)
        res[threads] = (used, _oracle_rate(po, head, steps), _oracle_rate(po, c1, c1_steps))
    po.set_threads(ncores)
    used, (rate, el), (rate64, el64) = res[ncores]
    _, (rate1, el1), (rate64_1, _) = res[1]
    return {"value": rate, "unit": "cell-updates/s", "cores": used, "kind": "port",
            "sample": "oracle/impec_oracle.c, %d steps of the headline workload at %d^2 after 1 untimed step "
                      "(same physics and tolerances), %d OpenMP threads, %.1f s" % (steps, n, used, el),
            "cpu_model": _cpu_model(), "nproc": ncores,
            "single_core_value": rate1, "single_core_s": el1,
            "c1_64": {"workload": "64^2, M=10, no dispersion, %d steps" % c1_steps,
                      "all_cores": rate64, "single_core": rate64_1}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=4096)
    ap.add_argument("--precision", choices=["fp64", "mixed", "mixed32"], default="fp64")
    ap.add_argument("--guess", choices=["zero", "previous"], default="zero")
    ap.add_argument("--mg-rows", type=int, default=0,
                    help="rows per wavefront chunk of the finest multigrid level (0: automatic)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-modes", action="store_true",
                    help="skip the plain / mixed / warm-start comparison on the lognormal-K workload")
    ap.add_argument("--no-roofline", action="store_true")
    ap.add_argument("--no-rotated", action="store_true", help="skip the rotated-grid line")
    ap.add_argument("--no-late-time", action="store_true",
                    help="skip the late-time line (the headline workload from a developed, fingered front)")
    ap.add_argument("--no-ensemble", action="store_true",
                    help="skip the batched-ensemble workload (256 realisations of mixed sizes)")
    ap.add_argument("--ensemble-members", type=int, default=256)
    ap.add_argument("--ensemble-steps", type=int, default=10)
    ap.add_argument("--ensemble-warmup", type=int, default=2)
    ap.add_argument("--ensemble-eager", action="store_true",
                    help="launch the ensemble's passes directly instead of replaying captured graphs")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run one plain 1 GiB device copy (reference kernel for HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: the IMPEC path runs on one GPU; --gpus %d is not supported" % args.gpus, file=sys.stderr)
        return 2

    from impec2d_amd import capi
    if args.calibrate_copy:
        import ctypes
        ms = ctypes.c_double(0)
        capi.load().impec_copy_calibration(1 << 30, 3, ctypes.byref(ms))
        print("copy calibration: 1 GiB in %.3f ms (%.0f GB/s read+write)" % (ms.value, 2 * 1.073741824 / ms.value * 1e3),
              file=sys.stderr)
    n = args.size
    ctx = capi.Context(workload_params(capi, n, args.precision, args.guess, args.mg_rows))
    ctx.set_source(one_cell_source(n))
    t = 0.0
    for _ in range(args.warmup):
        t += ctx.step(t)
    ctx.synchronize()
    s0 = ctx.stats()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        t += ctx.step(t)
    ctx.synchronize()
    el = time.perf_counter() - t0
    s1 = ctx.stats()
    vcyc = (s1.total_its - s0.total_its) / max(args.steps, 1)
    value = n * n * args.steps / el

    out = {
        "metric": "IMPEC cell-updates/s at 4096^2 (res 1e-10); V-cycles/step plain vs learned",
        "value": value, "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * el / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64" if args.precision == "fp64" else "f32/f64",
        "data": "synthetic (c = 0 at t = 0, one-cell wells)",
        "config": {"workload": "quarter-five-spot %d^2, M=10, Peaceman CN dispersion, nu=1 5-point, "
                               "MUSCL parallel splitting, CFL 0.5, rel-res 1e-10, %s, %s initial guess"
                               % (n, args.precision, args.guess),
                   "N": n, "vcycles_per_step": vcyc,
                   "parabolic_its_last": s1.last_parabolic_its, "pressure_relres_last": s1.last_relres,
                   "parabolic_relres_last": s1.last_parabolic_relres},
    }

    if not args.no_roofline:
        from impec2d_amd import roofline
        out["roofline"] = roofline.measure(ctx, n, t, fp32_cycle=args.precision in ("mixed", "mixed32"))
        out["roofline_by_kernel"] = roofline.measure.last_table
        out["step_roofline"] = roofline.step_summary(n, out["ms_per_step"])
    ctx.close()

    # the secondary workloads: a failure is reported in the line, never loses it
    def extra(key, fn):
        print("bench.py: %s ..." % key, file=sys.stderr, flush=True)
        try:
            out[key] = fn()
        except Exception as e:   # noqa: BLE001
            out[key] = {"error": "%s: %s" % (type(e).__name__, e)}
            print("bench.py: %s failed: %s" % (key, e), file=sys.stderr, flush=True)

    if not args.no_rotated:
        extra("rotated_grid", lambda: rotated(capi, n))
    if not args.no_late_time:
        extra("late_time", lambda: late_time(capi, n))
    if not args.no_modes:
        extra("modes", lambda: modes(capi, n))
    if not args.no_ensemble:
        extra("ensemble", lambda: ensemble(capi, members=args.ensemble_members, warmup=args.ensemble_warmup,
                                           steps=args.ensemble_steps, graphs=not args.ensemble_eager))
    if not args.no_cpu_baseline:
        extra("cpu_baseline", cpu_baseline)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
