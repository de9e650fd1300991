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
    # y = A x for the Crank-Nicolson operator, coefficients rebuilt from four
    # face-tensor arrays, fused with a BiCGStab dot product:
    # read x, d11, d12, d22, d21, dot operand (6 x 8) + write y (8)
    "cn_apply": dict(rocprof="k_cn_apply", bytes_per_cell=56),
    # one colour of red-black Gauss-Seidel on the fine level: read w (8), the
    # other colour's x (4), this colour's b (4), write this colour's x (4)
    "mg_smooth_l0": dict(rocprof="k_smooth<false, true>", bytes_per_cell=20),
    # x-sweep / y-sweep of MUSCL-Hancock: read c and one velocity component, write c
    "sweep_x": dict(rocprof="k_sweep_x", bytes_per_cell=24),
    "sweep_y": dict(rocprof="k_sweep_y", bytes_per_cell=24),
    # CG matrix-vector product with dot product: read d, w, write q
    "cg_matvec": dict(rocprof="k_cg_matvec", bytes_per_cell=24),
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


def measure(ctx, n, t, steps=2):
    """Runs `steps` profiled steps on ctx and returns the roofline object of
    the kernel class with the largest total time."""
    ctx.profile_reset()
    ctx.profile(True)
    for _ in range(steps):
        t += ctx.step(t)
    ctx.synchronize()
    ctx.profile(False)
    best = None
    for name in KERNELS:
        launches, ms = ctx.profile_read(name)
        if launches and (best is None or ms > best[2]):
            best = (name, launches, ms)
    if best is None:
        return None
    name, launches, ms = best
    avg_s = ms * 1e-3 / launches
    bytes_per_launch = KERNELS[name]["bytes_per_cell"] * float(n) * n
    achieved = bytes_per_launch / avg_s / 1e9
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": committed_traffic(name, n),
            "kernel": name, "rocprof_name": KERNELS[name]["rocprof"],
            "avg_launch_us": avg_s * 1e6, "launches": launches,
            "algorithmic_bytes_per_launch": bytes_per_launch}
