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


CPP_CALL_SITE = r'''
#include <vector>
#include "mhd_mi355x.hpp"

// Stand-in with the member names of the reference's Grid (Grid.hpp:9-26).
struct Grid {
  int nx, ny, ng = 2;
  double cell_width, cell_height, resistivity = 0.0;
  std::vector<double> rho, mx, my, E, bxf, byf, bxc, byc, Ez, vx, vy, p;
  Grid(int nx_, int ny_) : nx(nx_), ny(ny_), cell_width(0.1), cell_height(0.1) {
    const size_t n = size_t(nx + 4) * (ny + 4);
    for (auto* v : {&rho, &mx, &my, &E, &bxf, &byf, &bxc, &byc, &Ez, &vx, &vy, &p}) v->assign(n, 0.0);
  }
};

int main() {
  Grid grid(64, 64);
  try {
    mhd_mi355x::Context ctx(grid);
    double dt = mhd_mi355x::CFL_condition_checker(grid, ctx);
    mhd_mi355x::generate_A_matrix(ctx, dt * grid.resistivity * 0.2928932188134524);
    mhd_mi355x::IMEX(grid, ctx, dt);
    std::vector<double> d[6];
    for (auto& v : d) v.assign(grid.rho.size(), 0.0);
    mhd_mi355x::explicit_rhs(grid, ctx, d[0], d[1], d[2], d[3], d[4], d[5]);
    mhd_mi355x::resistive_B_term(grid, ctx, d[4], d[5], true);
    auto dv = mhd_mi355x::calculate_current_div_B(grid, ctx);
    (void)dv;
    mhd_mi355x::write_data_to_binary_file(grid, ctx, 0, 20, dt);
    int steps = mhd_mi355x::integrate(grid, ctx, 0.1, 0.0, 20);
    (void)steps;
  } catch (const mhd_mi355x::Error& e) {
    return e.code == MHD_EHIP ? 3 : 1;
  }
  return 0;
}
'''


def test_cpp_adapter_compiles(tmp_path):
    """A reference-style C++ call site compiles against include/mhd_mi355x.hpp
    and links libmhd_mi355x.so (it is not run: that needs a GPU)."""
    import shutil
    import subprocess
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    L.load()
    src = tmp_path / "call_site.cpp"
    src.write_text(CPP_CALL_SITE)
    pkg = os.path.dirname(L.LIB_PATH)
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), str(src),
                        "-o", str(tmp_path / "call_site"), os.path.join(pkg, "libmhd_mi355x.so"),
                        f"-Wl,-rpath,{pkg}"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


REFERENCE_INCLUDE = "/root/reference/include"


def test_cpp_adapter_compiles_on_reference_grid(tmp_path):
    """tests/cpp/reference_grid_call_site.cpp compiles against the reference's
    own Grid (include/Grid.hpp) where the reference tree is present, with
    -Wall -Werror: the adapter's templates take the real struct (the GPU test
    test_gpu_reference_grid.py runs the binary tests/cpp/Makefile builds)."""
    import shutil
    import subprocess
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    if not os.path.exists(os.path.join(REFERENCE_INCLUDE, "Grid.hpp")):
        pytest.skip("reference tree not present")
    L.load()
    pkg = os.path.dirname(L.LIB_PATH)
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I", REFERENCE_INCLUDE,
                        "-I", os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "cpp", "reference_grid_call_site.cpp"),
                        "-o", str(tmp_path / "call_site"), os.path.join(pkg, "libmhd_mi355x.so"),
                        f"-Wl,-rpath,{pkg}"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
