"""The C-ABI library loads and exports every entry point include/*.h declares
(no compute call is made without a GPU), and the product path has no CPU
fallback: without a device, mhd_create fails with MHD_EHIP."""
import ctypes as C
import glob
import os
import re

import pytest

import magnetohydrodynamics_simulator_amd._lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[\w\s\*]+?\b(mhd_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_header_declares_the_boundary():
    names = declared_symbols()
    for required in ("mhd_create", "mhd_destroy", "mhd_upload", "mhd_download", "mhd_imex_step",
                     "mhd_explicit_rhs", "mhd_resistive_term", "mhd_cg_solve", "mhd_cfl_dt", "mhd_divb",
                     "mhd_build_operator", "mhd_run", "mhd_last_error"):
        assert required in names


def test_library_exports_every_declared_symbol():
    lib = L.load()
    missing = [n for n in sorted(declared_symbols()) if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding covers the same set
    assert declared_symbols() <= set(L.SIGNATURES)


def test_default_options_match_the_reference_constants():
    lib = L.load()
    o = L.Opts()
    assert lib.mhd_default_opts(C.byref(o)) == 0
    assert o.gamma == 5.0 / 3.0 and o.cfl == 0.45 and o.cg_tol == 2.220446049250313e-16
    assert o.cg_max_iter == 0 and o.cg_fail_is_error == 0


def test_argument_errors_are_reported():
    lib = L.load()
    h = C.c_void_p()
    assert lib.mhd_create(2, 64, 0.1, 0.1, 0.0, None, C.byref(h)) == L.MHD_EINVAL
    assert b"nx and ny" in lib.mhd_last_error(None)
    assert lib.mhd_create(64, 64, -1.0, 0.1, 0.0, None, C.byref(h)) == L.MHD_EINVAL
    assert lib.mhd_create(64, 64, 0.1, 0.1, -1.0, None, C.byref(h)) == L.MHD_EINVAL


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_no_cpu_fallback_without_a_device():
    lib = L.load()
    h = C.c_void_p()
    rc = lib.mhd_create(64, 64, 0.1, 0.1, 0.0, None, C.byref(h))
    assert rc == L.MHD_EHIP and not h.value
    with pytest.raises(L.MhdError):
        from magnetohydrodynamics_simulator_amd import Context
        Context(64, 64, 0.1, 0.1)
