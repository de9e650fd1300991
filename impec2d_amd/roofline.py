"""Live HBM-roofline figure for the dominant kernel of an IMPEC step.

Each entry of KERNELS is a kernel class the library brackets with HIP events
on its own stream when profiling is enabled (impec_profile_enable), the
substring rocprofv3 lists it under, and its algorithmic bytes per cell of the
fine grid: compulsory traffic, every operand read once and every result
written once (FP64 = 8 B).
"""
import os
import re

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E

KERNELS = {
    # BiCGStab p-launch (k_cn_pstep): s = r - alpha v and t = A (s / D) again,
    # r' = s - omega t, x += alpha p/D + omega s/D, p' = r' + beta (p - omega v),
    # v' = A (p' / D), r0.v', r'.r'; read r, v, p, x (4 x 8), the shadow residual
    # r0 and 1/D (FP32, 2 x 4) and the face velocities u, v (2 x 8; the face
    # tensors are formed in registers), write r', x, p', v' (4 x 8)
    "cn_apply_p": dict(rocprof="k_cn_pstep", bytes_per_cell=88),
    # BiCGStab s-launch (k_cn_stream<1>): s = r - alpha v, t = A (s / D), t.s, t.t,
    # r0.t, r0.s, s.s; read r, v (2 x 8), r0 and 1/D (FP32, 2 x 4) and u, v
    # (2 x 8); no vector written
    "cn_apply_s": dict(rocprof="k_cn_stream<1>", bytes_per_cell=40),
    # zero-guess start (k_cn_start): b, 1/D, r0 = b in FP32 and v = A (b / D);
    # read c, u, v (3 x 8), write b (8), 1/D and r0 (2 x 4), v (8)
    "cn_start": dict(rocprof="k_cn_start", bytes_per_cell=48),
    # streaming V-cycle, finest level, down leg: read w, r (2 x 8), write x (8)
    # and the coarse right-hand side (8 / 4); FP32 cycle: w 4, r 8, x 4, b_c 1
    "mg_down_l0": dict(rocprof="k_mg_down<true, false", bytes_per_cell=34, bytes_per_cell_fp32=17),
    # up leg: read x, w, r (3 x 8) and the coarse solution (8 / 4), write x (8);
    # FP32 cycle: x 4, w 4, r 8, x_c 1, x 4
    "mg_up_l0": dict(rocprof="k_mg_up<true, false", bytes_per_cell=34, bytes_per_cell_fp32=21),
    # CG direction + matvec: read z, d_old, two face-mobility arrays (4 x 8),
    # write d, q (2 x 8); the FP32 cycle's z is 4 B
    "cg_dir_matvec": dict(rocprof="k_dir_mv", bytes_per_cell=48, bytes_per_cell_fp32=44),
    # CG update: read p, d, r, q (4 x 8), write p, r (2 x 8)
    "cg_update": dict(rocprof="k_cg_update2", bytes_per_cell=48),
    # symmetric parallel splitting in one pass (k_advect_parallel): read c, u, v
    # (3 x 8), write the averaged c (8)
    "advect_parallel": dict(rocprof="k_advect_parallel", bytes_per_cell=32),
    # MUSCL-Hancock x / y sweep: read c and one velocity component, write c
    "sweep_x": dict(rocprof="k_sweep_x", bytes_per_cell=24),
    "sweep_y": dict(rocprof="k_sweep_y", bytes_per_cell=24),
    # the smaller passes of a step (counted in the whole-step figure):
    # dispersion: the true-residual pass (read x, b, u, v) and the last
    # iteration's x update (read x, p, r, v and the FP32 1/D, write x)
    "cn_check": dict(rocprof="k_cn_stream<3>", bytes_per_cell=32),
    "cn_xfinal": dict(rocprof="k_cn_xfinal", bytes_per_cell=44),
    # the two at once (the last iteration's x update inside the true-residual
    # pass): read b, x, r, p, v (5 x 8), the FP32 1/D, u and v (2 x 8), write
    # the updated x (8)
    "cn_check_x": dict(rocprof="k_cn_stream<4>", bytes_per_cell=68),
    # velocities (read p and w, write u and v), CFL (read u, v), mobility
    # (read c, write w), the zero-guess CG start (read q, write p and r)
    "velocity": dict(rocprof="k_velocity", bytes_per_cell=32),
    "cfl": dict(rocprof="k_cfl", bytes_per_cell=16),
    "mobility": dict(rocprof="k_mobility", bytes_per_cell=16),
    "cg_init": dict(rocprof="k_cg_init_zero", bytes_per_cell=24),
}

_HERE = os.path.dirname(os.path.abspath(__file__))
_PROFILES = os.path.join(os.path.dirname(_HERE), "profiles")


def committed_traffic(name, n):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE and
    WRITE_SIZE scaled by a plain copy), from profiles/hbm_traffic_<name>_<n>.txt."""
    path = os.path.join(_PROFILES, "hbm_traffic_%s_%d.txt" % (name, n))
    if not os.path.exists(path):
        return None
    with open(path) as f:
        m = re.search(r"^TRAFFIC_BYTES_PER_LAUNCH\s+([0-9.eE+]+)", f.read(), re.M)
    return float(m.group(1)) if m else None


def step_summary(n, ms_per_step):
    """The whole step against the HBM roofline: the algorithmic bytes of every
    finest-level pass measured by the last measure() call (launches per step x
    bytes per cell x n^2) over the step time.  The coarse multigrid levels, the
    reductions' finishes, ghost fills and source cells are not counted (their
    bytes are a few percent of the step's), so this is a lower bound."""
    table, steps = getattr(measure, "last_table", None), getattr(measure, "last_steps", 0)
    if not table or not steps or not ms_per_step:
        return None
    b = sum(v["launches"] * v["bytes_per_cell"] for v in table.values()) / steps * float(n) * n
    kms = sum(v["launches"] * v["avg_us"] for v in table.values()) / steps / 1e3
    gbs = b / (ms_per_step * 1e-3) / 1e9
    return {"algorithmic_bytes_per_step": b, "bytes_per_cell": b / (float(n) * n), "achieved": gbs,
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
            "counted_kernel_ms_per_step": kms, "ms_per_step": ms_per_step}


def measure(ctx, n, t, steps=2, fp32_cycle=False):
    """Runs `steps` profiled steps on ctx and returns the roofline object of
    the kernel class with the largest total time (fp32_cycle: the
    mixed-precision V-cycle's byte counts)."""
    ctx.profile_reset()
    ctx.profile(True)
    for _ in range(steps):
        t += ctx.step(t)
    ctx.synchronize()
    ctx.profile(False)
    best = None
    table = {}
    for name in KERNELS:
        launches, ms = ctx.profile_read(name)
        if not launches:
            continue
        k = KERNELS[name]
        per_cell = k.get("bytes_per_cell_fp32", k["bytes_per_cell"]) if fp32_cycle else k["bytes_per_cell"]
        avg = ms * 1e-3 / launches
        gbs = per_cell * float(n) * n / avg / 1e9
        table[name] = {"launches": launches, "avg_us": round(avg * 1e6, 2), "GB/s": round(gbs, 1),
                       "frac": round(gbs / HBM_PEAK_GBS, 4), "bytes_per_cell": per_cell}
        if best is None or ms > best[2]:
            best = (name, launches, ms)
    measure.last_table = table
    measure.last_steps = steps
    if best is None:
        return None
    name, launches, ms = best
    avg_s = ms * 1e-3 / launches
    k = KERNELS[name]
    per_cell = k.get("bytes_per_cell_fp32", k["bytes_per_cell"]) if fp32_cycle else k["bytes_per_cell"]
    bytes_per_launch = per_cell * float(n) * n
    achieved = bytes_per_launch / avg_s / 1e9
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": committed_traffic(name, n),
            "kernel": name, "rocprof_name": k["rocprof"],
            "avg_launch_us": avg_s * 1e6, "launches": launches,
            "algorithmic_bytes_per_launch": bytes_per_launch}
