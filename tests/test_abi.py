"""The C ABI library loads and exports every entry point of include/impec_hip.h,
and the ctypes mirrors of its structs match the C layout.  No GPU needed."""
import os
import re
import subprocess
import tempfile

import pytest

from impec2d_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "impec_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(impec_\w+)\s*\(", text, re.M)))


def test_header_lists_entry_points():
    names = declared_functions()
    assert "impec_create" in names and "impec_step" in names and len(names) >= 35


def test_library_exports_every_declared_symbol():
    lib = capi.load()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", capi.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (impec_\w+)", out))
    assert set(declared_functions()) <= exported


def test_ctypes_signatures_cover_header():
    assert set(declared_functions()) == set(capi.SIGNATURES)


def test_struct_layout_matches_c():
    import ctypes as C
    src = r'''
#include <stdio.h>
#include <stddef.h>
#include "impec_hip.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu\n", sizeof(impec_params), offsetof(impec_params, x_lb),
         offsetof(impec_params, inner_tol), sizeof(impec_stats), offsetof(impec_stats, last_relres),
         offsetof(impec_stats, total_mass_outflux));
  return 0;
}
'''
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        exe = os.path.join(d, "t")
        open(c, "w").write(src)
        subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe])
        vals = [int(x) for x in subprocess.check_output([exe]).split()]
    mine = [C.sizeof(capi.Params), capi.Params.x_lb.offset, capi.Params.inner_tol.offset,
            C.sizeof(capi.Stats), capi.Stats.last_relres.offset, capi.Stats.total_mass_outflux.offset]
    assert vals == mine


def test_params_default_matches_example_configs():
    p = capi.default_params()
    assert p.num_ghostcells == 2 and p.max_iterations == 10000
    assert abs(p.nu - 2 / 3.0) < 1e-15 and p.cfl == 0.5 and p.mu_res == 41.0


def test_no_gpu_fails_loudly():
    """Without a device the product path refuses to run (no CPU fallback)."""
    lib = capi.load()
    if lib.impec_device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(capi.ImpecError) as e:
        capi.Context(capi.default_params())
    assert e.value.code == capi.IMPEC_ENODEV


def test_invalid_params_rejected():
    import ctypes as C
    lib = capi.load()
    h = C.c_void_p()
    for kw in [dict(num_ghostcells=3), dict(N1=1), dict(nu=1.5), dict(source_type=2), dict(dim_splitting=7),
               dict(mg_variant=3), dict(precision=3), dict(N1=65536, N2=32768)]:
        p = capi.default_params(**kw)
        assert lib.impec_create(C.byref(p), C.byref(h)) == capi.IMPEC_EINVAL


def test_ensemble_member_validation():
    """impec_ensemble_create checks every member before touching a device:
    members outside the batched path (nu < 1, dispersion, other splittings,
    test mode, learned guess, mixed precisions or schemes, the FP32-vector CG)
    are refused, an
    acceptable set without a device fails with IMPEC_ENODEV."""
    import ctypes as C
    lib = capi.load()
    h = C.c_void_p()
    good = dict(nu=1.0, dim_splitting=4)

    def create(*members):
        arr = (capi.Params * len(members))(*members)
        return lib.impec_ensemble_create(arr, len(members), C.byref(h))

    base = capi.default_params(**good)
    for kw in [dict(nu=2 / 3.0), dict(diffusion=1), dict(dim_splitting=0), dict(is_test=1),
               dict(initial_guess=capi.GUESS_LEARNED), dict(mg_variant=1)]:
        assert create(base, capi.default_params(**dict(good, **kw))) == capi.IMPEC_EINVAL, kw
    assert create(base, capi.default_params(**dict(good, precision=capi.PREC_MIXED))) == capi.IMPEC_EINVAL
    # the FP32-vector CG runs on single contexts only, even when every member asks for it
    m32 = capi.default_params(**dict(good, precision=capi.PREC_MIXED32))
    assert create(m32, m32) == capi.IMPEC_EINVAL
    assert create(base, capi.default_params(**dict(good, hyper_method=1))) == capi.IMPEC_EINVAL
    assert lib.impec_ensemble_create(None, 0, C.byref(h)) == capi.IMPEC_EINVAL
    if lib.impec_device_count() == 0:
        assert create(base, capi.default_params(**dict(good, N1=128, N2=96))) == capi.IMPEC_ENODEV
