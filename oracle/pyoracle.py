"""ctypes bindings for the CPU oracle (TEST INFRASTRUCTURE).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module; the product package never does.  The oracle restates the
reference's header-only solver (include/Grid.hpp, MUSCL.hpp, HLL.hpp,
solvers.hpp, diagnostics.hpp, benchmarks.hpp and the driver loop in
src/integrator_main.cpp) in plain C; see oracle/mhd_oracle.c for the per-
function citations and what pins each function.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
FIELDS = ("rho", "mx", "my", "E", "bxf", "byf", "vx", "vy", "p", "bxc", "byc", "Ez")
CONSERVED = ("rho", "mx", "my", "E", "bxf", "byf")

_dp = C.POINTER(C.c_double)


class StepStats(C.Structure):
    _fields_ = [("cg_iters", C.c_int * 2), ("cg_spmv", C.c_int * 2), ("cg_error", C.c_double * 2)]


def build(fast_native: bool = False, out_dir: str | None = None) -> None:
    """Compile the oracle libraries with oracle/Makefile (gcc, seconds)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True, stdout=subprocess.DEVNULL)
    if fast_native:
        out = os.path.join(out_dir or HERE, "libmhd_oracle_native.so")
        subprocess.run(["gcc", "-std=c11", "-O3", "-ffast-math", "-march=native", "-fPIC", "-shared",
                        "-o", out, os.path.join(HERE, "mhd_oracle.c"), "-lm"], check=True)


def _bind(lib: C.CDLL) -> C.CDLL:
    vp = C.c_void_p
    sig = {
        "og_grid_create": (vp, [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double]),
        "og_grid_destroy": (None, [vp]),
        "og_field": (_dp, [vp, C.c_int]),
        "og_grid_copy": (None, [vp, vp]),
        "og_grid_set_eta": (None, [vp, C.c_double]),
        "og_fill_ghost_periodic": (None, [vp]),
        "og_fill_ghost": (None, [vp]),
        "og_grid_set_bc": (None, [vp, C.c_int, C.c_int]),
        "og_compute_cell_b": (None, [vp]),
        "og_compute_primitives": (None, [vp]),
        "og_ws_create": (vp, [C.c_int, C.c_int]),
        "og_ws_destroy": (None, [vp]),
        "og_ws_set_cg": (None, [vp, C.c_double, C.c_longlong]),
        "og_ws_rv": (_dp, [vp, C.c_int]),
        "og_ws_flux": (_dp, [vp, C.c_int]),
        "og_minmod": (C.c_double, [C.c_double, C.c_double]),
        "og_fast_speed": (C.c_double, [_dp, C.c_double]),
        "og_hll_x": (None, [_dp, _dp, _dp]),
        "og_hll_y": (None, [_dp, _dp, _dp]),
        "og_muscl_all": (None, [vp, vp]),
        "og_explicit_rhs": (None, [vp, vp, C.POINTER(_dp)]),
        "og_generate_A_matrix": (None, [vp, vp, C.c_double]),
        "og_cg_solve": (None, [vp, _dp, _dp, C.POINTER(C.c_int), C.POINTER(C.c_int), _dp]),
        "og_resistive_term": (None, [vp, vp, C.c_int, _dp, _dp, C.POINTER(C.c_int), C.POINTER(C.c_int), _dp]),
        "og_imex": (None, [vp, vp, C.c_double, C.POINTER(StepStats)]),
        "og_cfl": (C.c_double, [vp, C.c_double]),
        "og_divb": (None, [vp, _dp, _dp]),
        "og_ic_orszag_tang": (None, [vp, vp]),
        "og_ic_alfven": (None, [vp]),
        "og_ic_gaussian": (None, [vp]),
        "og_ic_acoustic": (None, [vp]),
        "og_ic_sausage": (None, [vp]),
        "og_ic_kink": (None, [vp]),
        "og_ic_decaying_sine": (None, [vp, vp]),
        "og_run": (C.c_int, [vp, vp, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                             _dp, C.POINTER(C.c_int), _dp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_LIBS: dict[str, C.CDLL] = {}


def load(variant: str = "strict", path: str | None = None) -> C.CDLL:
    """variant 'strict' (parity checker) or 'fast' (reference flags)."""
    key = path or variant
    if key not in _LIBS:
        if path is None:
            name = {"strict": "libmhd_oracle.so", "fast": "libmhd_oracle_fast.so"}[variant]
            path = os.path.join(HERE, name)
        if not os.path.exists(path):
            build()
        _LIBS[key] = _bind(C.CDLL(path))
    return _LIBS[key]


def _ptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


class Sim:
    """One reference Grid plus its integrator workspace (oracle side)."""

    def __init__(self, nx: int, ny: int, dx: float, dy: float, eta: float = 0.0,
                 variant: str = "strict", lib: C.CDLL | None = None, bc: tuple = ("periodic", "periodic")):
        """bc: boundary of the x and y axes, "periodic" (the reference's) or
        "outflow" (the zero-gradient ghost rule of og_grid_set_bc)."""
        self.lib = lib or load(variant)
        self.nx, self.ny = nx, ny
        self.nxt, self.nyt = nx + 4, ny + 4
        self.dx, self.dy, self.eta = dx, dy, eta
        codes = {"periodic": 0, "outflow": 1}
        bcx, bcy = (codes[b] for b in bc)
        self.g = self.lib.og_grid_create(nx, ny, dx, dy, eta)
        self.lib.og_grid_set_bc(self.g, bcx, bcy)
        self.w = self.lib.og_ws_create(nx, ny)

    def close(self):
        if self.g:
            self.lib.og_grid_destroy(self.g)
            self.lib.og_ws_destroy(self.w)
            self.g = self.w = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def field(self, name: str) -> np.ndarray:
        """Live (nyt, nxt) view of a padded Grid array."""
        p = self.lib.og_field(self.g, FIELDS.index(name))
        return np.ctypeslib.as_array(p, shape=(self.nyt, self.nxt))

    def interior(self, name: str) -> np.ndarray:
        return self.field(name)[2:-2, 2:-2].copy()

    def set_conserved(self, arrays: dict) -> None:
        for k, v in arrays.items():
            self.field(k)[...] = v

    def set_cg(self, tol: float, max_iter: int = -1) -> None:
        self.lib.og_ws_set_cg(self.w, tol, max_iter)

    def fill_ghost(self):
        self.lib.og_fill_ghost(self.g)

    def primitives(self):
        self.lib.og_compute_primitives(self.g)

    def cfl(self, cfl: float = 0.45) -> float:
        return self.lib.og_cfl(self.g, cfl)

    def divb(self):
        a, b = C.c_double(), C.c_double()
        self.lib.og_divb(self.g, C.byref(a), C.byref(b))
        return a.value, b.value

    def build_operator(self, alpha: float):
        self.lib.og_generate_A_matrix(self.g, self.w, alpha)

    def explicit_rhs(self):
        outs = [np.zeros((self.nyt, self.nxt)) for _ in range(6)]
        arr = (_dp * 6)(*[_ptr(o) for o in outs])
        self.lib.og_explicit_rhs(self.g, self.w, arr)
        return outs

    def cg_solve(self, b: np.ndarray):
        b = np.ascontiguousarray(b, dtype=np.float64).ravel()
        x = np.zeros_like(b)
        it, sp, err = C.c_int(), C.c_int(), C.c_double()
        self.lib.og_cg_solve(self.w, _ptr(b), _ptr(x), C.byref(it), C.byref(sp), C.byref(err))
        return x, it.value, sp.value, err.value

    def resistive_term(self, implicit: bool):
        bx = np.zeros((self.nyt, self.nxt)); by = np.zeros((self.nyt, self.nxt))
        it, sp, err = C.c_int(), C.c_int(), C.c_double()
        self.lib.og_resistive_term(self.g, self.w, int(implicit), _ptr(bx), _ptr(by),
                                   C.byref(it), C.byref(sp), C.byref(err))
        return bx, by, it.value

    def imex(self, dt: float) -> StepStats:
        st = StepStats()
        self.lib.og_imex(self.g, self.w, dt, C.byref(st))
        return st

    def ic(self, name: str):
        f = {"orszag_tang": lambda: self.lib.og_ic_orszag_tang(self.g, self.w),
             "alfven": lambda: self.lib.og_ic_alfven(self.g),
             "gaussian": lambda: self.lib.og_ic_gaussian(self.g),
             "acoustic": lambda: self.lib.og_ic_acoustic(self.g),
             "sausage": lambda: self.lib.og_ic_sausage(self.g),
             "kink": lambda: self.lib.og_ic_kink(self.g),
             "decaying_sine": lambda: self.lib.og_ic_decaying_sine(self.g, self.w)}[name]
        f()

    def run(self, t_end: float, cfl: float = 0.45, dt_override: float = 0.0,
            max_steps: int = -1, snapshot_every: int = 20):
        cap = max_steps if max_steps > 0 else 1_000_000
        cap = min(cap, 1_000_000)
        dts = np.zeros(cap)
        its = np.zeros(2 * cap, dtype=np.int32)
        t = C.c_double()
        n = self.lib.og_run(self.g, self.w, t_end, dt_override, cfl, max_steps, snapshot_every,
                            _ptr(dts), its.ctypes.data_as(C.POINTER(C.c_int)), C.byref(t))
        return n, dts[:n].copy(), its[: 2 * n].reshape(n, 2).copy(), t.value


class RefPartial:
    """The reference's own Grid/MUSCL/HLL/diagnostics headers (oracle/_ref)."""

    def __init__(self, path: str | None = None):
        path = path or os.path.join(HERE, "_ref", "libref_partial.so")
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        lib = C.CDLL(path)
        pp = C.POINTER(_dp)
        lib.ref_fill_and_primitives.argtypes = [C.c_int, C.c_int, pp]
        lib.ref_muscl_all.argtypes = [C.c_int, C.c_int, pp, pp]
        lib.ref_minmod.restype = C.c_double
        lib.ref_minmod.argtypes = [C.c_double, C.c_double]
        lib.ref_fast_speed.restype = C.c_double
        lib.ref_fast_speed.argtypes = [_dp, C.c_double]
        lib.ref_hll_x.argtypes = [_dp, _dp, _dp]
        lib.ref_hll_y.argtypes = [_dp, _dp, _dp]
        lib.ref_cfl.restype = C.c_double
        lib.ref_cfl.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, pp]
        lib.ref_divb.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, pp, _dp, _dp]
        self.lib = lib

    @staticmethod
    def ptrs(arrays):
        return (_dp * len(arrays))(*[(_ptr(a) if a is not None else None) for a in arrays])


def padded_pointers(arrays):
    return (_dp * len(arrays))(*[_ptr(a) for a in arrays])
