"""The C-ABI libraries load on a machine without a GPU and export every
entry point include/n4_capi.h declares; the compatibility library exports the
reference's C++ API names (include/cudaN4.h, include/reducer.h).  No compute
call is made here."""
import ctypes as C
import os
import re
import subprocess

import pytest

import n4amd

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "n4_capi.h")


def declared_functions():
    text = open(HEADER).read()
    decl = re.findall(r"^(?:n4_status|const char \*)\s*(n4_[a-z0-9_]+)\(", text, flags=re.M)
    return sorted(set(decl))


def test_header_and_binding_agree():
    assert set(declared_functions()) == set(n4amd.EXPORTS)


@pytest.mark.parametrize("path", [n4amd.LIB_PATH, n4amd.EMU_PATH])
def test_library_exports_every_declared_symbol(path):
    lib = C.CDLL(path)
    for name in declared_functions():
        assert hasattr(lib, name), f"{os.path.basename(path)} lacks {name}"


def test_host_only_calls_without_gpu():
    lib = C.CDLL(n4amd.LIB_PATH)
    lib.n4_status_string.restype = C.c_char_p
    assert lib.n4_status_string(0) == b"ok"
    p = n4amd.Params()
    assert lib.n4_params_default(C.byref(p)) == 0
    assert p.number_of_histogram_bins == 200 and abs(p.bias_field_fwhm - 0.15) < 1e-15
    nc = (C.c_uint32 * 3)()
    assert lib.n4_final_ncpt(C.byref(p), nc) == 0
    assert list(nc) == [11, 11, 11]  # (4-3)*2^3+3, main.cpp:83-92
    p.ncpt[0] = 3
    assert lib.n4_final_ncpt(C.byref(p), nc) == 1  # N4_ERR_INVALID_ARG


def test_compat_library_exports_reference_api():
    so = os.path.join(n4amd.BUILD_DIR, "libn4compat.so")
    out = subprocess.run(["nm", "-DC", "--defined-only", so], capture_output=True, text=True, check=True).stdout
    for name in ["N4(", "testN4(", "sharpenImage(", "testsharpenImage(", "calculateConvergenceMeasurement(",
                 "upsampleLattice3D_cpu(", "upsampleLattice3D_gpu(", "testUpsampleLattice3D_gpu(",
                 "EvaluateBspline3D(", "testEvaluateBspline3D(", "FitBspline3D(", "testFitBspline3D(",
                 "N4Data::assign(", "N4Data::clean(", "Reducer::reduce_sum_wrapper(",
                 "Reducer::reduce_max_wrapper(", "Reducer::reduce_min_wrapper("]:
        assert name in out, name
