"""Builds libimpec_hip.so in-tree for gfx950 with hipcc.

Run as ``python -m impec2d_amd.build`` or through __graft_entry__.build().
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libimpec_hip.so")

SOURCES = ["capi.hip", "k_field.hip", "k_pressure.hip", "k_transport.hip", "k_cn.hip", "k_cnn.hip", "ensemble.hip", "noalloc.hip"]

# FP64 paths must round like the reference's x86-64 build: no contraction of
# a*b+c into fused multiply-adds.  Packed FP32 instructions are disabled.
HIPCC_FLAGS = [
    "-O3", "--offload-arch=gfx950", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-value", "-Wno-unused-result",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose=False, jobs=8, out=LIB, defines=()):
    """Compiles every source to an object (in parallel) and links the .so.

    `defines` are extra -D options (the kernels' tuning constants, e.g.
    "IMPEC_SW_PF=2"); `out` is the library path (objects go beside it)."""
    objdir = os.path.join(os.path.dirname(os.path.abspath(out)), "_build")
    if out != LIB:
        objdir = os.path.join(os.path.dirname(os.path.abspath(out)), "_build_" + os.path.basename(out))
    os.makedirs(objdir, exist_ok=True)
    procs = []
    objs = []
    for src in SOURCES:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        objs.append(obj)
        cmd = [hipcc()] + HIPCC_FLAGS + ["-D" + d for d in defines] + ["-I", os.path.join(ROOT, "include"), "-c",
                                                                      os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        if len([p for _, p in procs if p.poll() is None]) >= jobs:
            procs[0][1].wait()
    failed = []
    for src, p in procs:
        log, _ = p.communicate()
        text = log.decode(errors="replace")
        # hipcc's host pass does not know the packed-fp32 feature: harmless
        lines = [l for l in text.splitlines() if "packed-fp32-ops" not in l and l.strip()]
        if p.returncode != 0:
            failed.append(src)
            print("\n".join(lines), file=sys.stderr)
        elif verbose and lines:
            print("\n".join(lines))
    if failed:
        raise RuntimeError("hipcc failed for: " + ", ".join(failed))
    tmp = out + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs
    subprocess.check_call(cmd)
    os.replace(tmp, out)
    return out


HOST_LIB = os.path.join(HERE, "libimpec_host.so")
HOST_RUN = os.path.join(HERE, "impec2d_run")
GXX_FLAGS = ["-std=c++17", "-O2", "-fPIC", "-Wall", "-Wextra", "-Wno-unused-parameter"]


HOST_SOURCES = ["impec_solvers.cpp", "impec_config.cpp", "impec_output.cpp", "impec_tests.cpp"]


def build_host(verbose=False):
    """The C++ solver classes, configuration reader and output path (host/) as libimpec_host.so and
    the impec2d_run driver, both linked against libimpec_hip.so (g++ only: no
    device code on this side)."""
    gxx = os.environ.get("CXX", "g++")
    inc = ["-I", os.path.join(ROOT, "include"), "-I", os.path.join(HERE, "host")]
    link = ["-L", HERE, "-limpec_hip", "-Wl,-rpath,$ORIGIN"]
    steps = [
        [gxx] + GXX_FLAGS + ["-shared"] + inc + [os.path.join(HERE, "host", f) for f in HOST_SOURCES] +
        ["-o", HOST_LIB + ".tmp"] + link,
        [gxx] + GXX_FLAGS + inc + [os.path.join(HERE, "host", "impec2d_run.cpp"), "-o", HOST_RUN + ".tmp",
                                   "-L", HERE, "-limpec_host"] + link,
    ]
    for i, cmd in enumerate(steps):
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.check_call(cmd)
        if i == 0:
            os.replace(HOST_LIB + ".tmp", HOST_LIB)
    os.replace(HOST_RUN + ".tmp", HOST_RUN)
    return HOST_LIB, HOST_RUN


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    print(build_host(verbose="-v" in sys.argv))
