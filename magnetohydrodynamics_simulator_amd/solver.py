"""Host-side mirror of the reference's step interface over the gfx950 C ABI.

The reference (RikibabManos/Magnetohydrodynamics_Simulator; citations relative
to its repository) is a header-only C++ program whose hot path is a set of
free functions over a padded ``Grid`` (include/Grid.hpp:6-62):
``IMEX`` (include/solvers.hpp:491), ``explicit_rhs`` (:393),
``resistive_B_term`` (:445), ``generate_A_matrix`` (:404),
``CFL_condition_checker`` and ``calculate_current_div_B``
(include/diagnostics.hpp:44,15).  This module keeps those names and argument
meanings; the Eigen solver and the scratch structs become a ``Context`` that
owns the device state.  All arrays at this boundary are padded (ny+4, nx+4)
float64, exactly the reference Grid's storage.
"""
from __future__ import annotations

import ctypes as C
import os
import math

import numpy as np

from . import _lib

GAMMA_ARS = (2.0 - math.sqrt(2.0)) / 2.0      # IMEX_Butcher_Tableus::gamma, solvers.hpp:41
PI_REF = 3.14159265359                         # integrator_main.cpp:36

_dp = C.POINTER(C.c_double)

# The reference's setups by name (the keys of setups.SETUPS) and their
# MHD_SETUP_* ids; benchmarks.hpp:419-451 selects the same seven.
SETUP_IDS = {"orszag_tang": 0, "decaying_sine": 1, "alfven": 2, "gaussian": 3, "acoustic": 4, "sausage": 5,
             "kink": 6}


def _setup_id(name: str) -> int:
    try:
        return SETUP_IDS[name]
    except KeyError:
        raise ValueError(f"unknown setup {name!r}; one of {sorted(SETUP_IDS)}") from None


def _ptr(a: np.ndarray | None):
    if a is None:
        return None
    if a.dtype != np.float64 or not a.flags.c_contiguous:
        raise TypeError("arrays must be C-contiguous float64")
    return a.ctypes.data_as(_dp)


class Grid:
    """Host copy of the reference Grid: 12 padded float64 arrays, ng = 2."""

    FIELDS = ("rho", "mx", "my", "E", "bxc", "byc", "bxf", "byf", "Ez", "vx", "vy", "p")

    def __init__(self, nx: int, ny: int, ng: int = 2):
        if ng != 2:
            raise ValueError("the MUSCL path uses ng = 2 (Grid.hpp:29)")
        self.nx, self.ny, self.ng = nx, ny, ng
        self.nxt, self.nyt = nx + 2 * ng, ny + 2 * ng
        self.cell_width = 2 * PI_REF / nx
        self.cell_height = 2 * PI_REF / ny
        self.resistivity = 0.0
        for f in self.FIELDS:
            setattr(self, f, np.zeros((self.nyt, self.nxt)))

    def indexC(self, i: int, j: int) -> int:
        return j * self.nxt + i

    def physical(self, name: str) -> np.ndarray:
        return getattr(self, name)[2:-2, 2:-2]

    def copy(self) -> "Grid":
        g = Grid(self.nx, self.ny, self.ng)
        g.cell_width, g.cell_height, g.resistivity = self.cell_width, self.cell_height, self.resistivity
        for f in self.FIELDS:
            setattr(g, f, getattr(self, f).copy())
        return g


class Context:
    """Device state + implicit operator of one run (replaces the reference's
    integrator_reserved_memory, ReconstructedValues, the two scratch Grids and
    the Eigen ConjugateGradient object)."""

    def __init__(self, nx: int, ny: int, dx: float, dy: float, eta: float = 0.0, *, cfl: float = 0.45,
                 cg_tol: float | None = None, cg_max_iter: int = 0, cg_fail_is_error: bool = False,
                 march_rows: int = 0, bc: tuple = ("periodic", "periodic")):
        """bc: boundary of the x and y axes, "periodic" (the reference's) or
        "outflow" (zero-gradient; MHD_BC_OUTFLOW of the C ABI)."""
        self.lib = _lib.load()
        o = _lib.Opts()
        _lib.check(self.lib.mhd_default_opts(C.byref(o)))
        o.cfl = cfl
        if cg_tol is not None:
            o.cg_tol = cg_tol
        o.cg_max_iter = cg_max_iter
        o.cg_fail_is_error = int(cg_fail_is_error)
        o.march_rows = march_rows
        try:
            o.bc_x, o.bc_y = (_lib.BC_NAMES[b] for b in bc)
        except (KeyError, ValueError, TypeError):
            raise ValueError(f"bc must be a pair of 'periodic' / 'outflow', got {bc!r}") from None
        self.bc = tuple(bc)
        h = C.c_void_p()
        _lib.check(self.lib.mhd_create(nx, ny, dx, dy, eta, C.byref(o), C.byref(h)), None)
        self.h = h
        self.nx, self.ny, self.dx, self.dy, self.eta = nx, ny, dx, dy, eta
        self.nxt, self.nyt = nx + 4, ny + 4
        self.opts = o

    @classmethod
    def for_grid(cls, g: Grid, **kw) -> "Context":
        return cls(g.nx, g.ny, g.cell_width, g.cell_height, g.resistivity, **kw)

    def close(self):
        if getattr(self, "h", None):
            self.lib.mhd_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        _lib.check(rc, self.h)

    def _padded(self):
        return np.zeros((self.nyt, self.nxt))

    # ---- state transfer ----
    def upload(self, rho, mx, my, E, bxf, byf):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (rho, mx, my, E, bxf, byf)]
        for a in arrs:
            if a.shape != (self.nyt, self.nxt):
                raise ValueError(f"expected padded shape {(self.nyt, self.nxt)}, got {a.shape}")
        self._check(self.lib.mhd_upload(self.h, *[_ptr(a) for a in arrs]))

    def upload_grid(self, g: Grid):
        self.upload(g.rho, g.mx, g.my, g.E, g.bxf, g.byf)

    def download(self, primitives: bool = False) -> dict:
        names = ["rho", "mx", "my", "E", "bxf", "byf"]
        out = {n: self._padded() for n in names}
        prim = ["vx", "vy", "p", "bxc", "byc"]
        if primitives:
            out.update({n: self._padded() for n in prim})
        ptrs = [_ptr(out[n]) for n in names] + [(_ptr(out[n]) if primitives else None) for n in prim]
        self._check(self.lib.mhd_download(self.h, *ptrs))
        return out

    def download_grid(self, g: Grid):
        d = self.download(primitives=True)
        for k, v in d.items():
            getattr(g, k)[...] = v

    # ---- the reference's hot-path entry points ----
    def CFL_condition_checker(self) -> float:
        dt = C.c_double()
        self._check(self.lib.mhd_cfl_dt(self.h, C.byref(dt)))
        return dt.value

    def generate_A_matrix(self, alpha: float):
        self._check(self.lib.mhd_build_operator(self.h, alpha))

    def IMEX(self, dt: float) -> _lib.StepStats:
        st = _lib.StepStats()
        self._check(self.lib.mhd_imex_step(self.h, dt, C.byref(st)))
        return st

    def explicit_rhs(self):
        outs = [self._padded() for _ in range(6)]
        self._check(self.lib.mhd_explicit_rhs(self.h, *[_ptr(o) for o in outs]))
        return outs

    def resistive_B_term(self, implicit: bool):
        bx, by = self._padded(), self._padded()
        it = C.c_int()
        self._check(self.lib.mhd_resistive_term(self.h, int(implicit), _ptr(bx), _ptr(by), C.byref(it)))
        return bx, by, it.value

    def cg_solve(self, b: np.ndarray):
        b = np.ascontiguousarray(b, dtype=np.float64).ravel()
        if b.size != self.nx * self.ny:
            raise ValueError("b must hold nx * ny corner values")
        x = np.zeros_like(b)
        it, err = C.c_int(), C.c_double()
        self._check(self.lib.mhd_cg_solve(self.h, _ptr(b), _ptr(x), C.byref(it), C.byref(err)))
        return x, it.value, err.value

    def calculate_current_div_B(self):
        rms, mx = C.c_double(), C.c_double()
        self._check(self.lib.mhd_divb(self.h, C.byref(rms), C.byref(mx)))
        return rms.value, mx.value

    def run(self, t_end: float, dt_override: float = 0.0, max_steps: int = -1):
        """The driver loop of src/integrator_main.cpp:52-112 (no snapshots)."""
        cap = max_steps if max_steps > 0 else 1_000_000
        dts = np.zeros(cap)
        its = np.zeros(2 * cap, dtype=np.int32)
        steps, t = C.c_int(), C.c_double()
        self._check(self.lib.mhd_run(self.h, t_end, dt_override, max_steps, C.byref(steps), C.byref(t),
                                     _ptr(dts), its.ctypes.data_as(C.POINTER(C.c_int))))
        n = steps.value
        return n, dts[:n].copy(), its[: 2 * n].reshape(n, 2).copy(), t.value

    def run_resident(self, t_end: float, dt_override: float = 0.0, max_steps: int = -1, snapshot_every: int = 0,
                     on_snapshot=None):
        """The same driver loop with its per-step decisions on the device
        (mhd_run_resident): identical results to run(), one host read-back
        per group of steps.  on_snapshot(ctx, frame, t, dt) runs before the
        first step and after every snapshot_every-th step; it may download
        or write the state."""
        cap = max_steps if max_steps > 0 else 1_000_000
        dts = np.zeros(cap)
        its = np.zeros(2 * cap, dtype=np.int32)
        steps, t = C.c_int(), C.c_double()
        raised = []

        def trampoline(_h, frame, tt, dt, _user):
            try:
                on_snapshot(self, frame, tt, dt)
                return 0
            except BaseException as e:   # re-raised below, after the C loop has returned
                raised.append(e)
                return _lib.MHD_EINVAL

        cb = _lib.SNAPSHOT_FN(trampoline) if on_snapshot is not None else None
        rc = self.lib.mhd_run_resident(self.h, t_end, dt_override, max_steps, snapshot_every,
                                       C.cast(cb, C.c_void_p) if cb is not None else None, None,
                                       C.byref(steps), C.byref(t), _ptr(dts), its.ctypes.data_as(C.POINTER(C.c_int)))
        if raised:
            raise raised[0]
        self._check(rc)
        n = steps.value
        return n, dts[:n].copy(), its[: 2 * n].reshape(n, 2).copy(), t.value

    def run_resident_files(self, t_end: float, snapshot_every: int, path_format: str, snapshot_rate: int | None = None,
                           dt_override: float = 0.0, max_steps: int = -1):
        """run_resident with the reference's snapshot files written while the
        loop runs on (mhd_run_resident_files): path_format holds one %d (the
        frame = steps taken), e.g. "output/snapshot_%06d.dat"; the header's
        snapshot rate is snapshot_rate (default snapshot_every).  Returns
        after the last file is written."""
        cap = max_steps if max_steps > 0 else 1_000_000
        dts = np.zeros(cap)
        its = np.zeros(2 * cap, dtype=np.int32)
        steps, t = C.c_int(), C.c_double()
        rate = snapshot_every if snapshot_rate is None else snapshot_rate
        self._check(self.lib.mhd_run_resident_files(self.h, t_end, dt_override, max_steps, snapshot_every,
                                                    os.fsencode(path_format), rate, C.byref(steps), C.byref(t),
                                                    _ptr(dts), its.ctypes.data_as(C.POINTER(C.c_int))))
        n = steps.value
        return n, dts[:n].copy(), its[: 2 * n].reshape(n, 2).copy(), t.value

    def resident_tune(self, group_steps: int = 16, spec_margin: int = 0, fixed_budget: int = -1,
                      graph: int = 0):
        """Group size, CG iteration margin, an optional fixed iteration budget
        and graph replay of run_resident (mhd_resident_tune; graph 1 / True:
        captured pairs of steps, 2: each group as one captured graph)."""
        self._check(self.lib.mhd_resident_tune(self.h, group_steps, spec_margin, fixed_budget, int(graph)))

    # ---- adjoint ----
    def _field_ptrs(self, arrs):
        keep = [np.ascontiguousarray(a, dtype=np.float64) for a in arrs]
        for a in keep:
            if a.shape != (self.nyt, self.nxt):
                raise ValueError(f"expected padded shape {(self.nyt, self.nxt)}, got {a.shape}")
        return keep, (_dp * 6)(*[_ptr(a) for a in keep])

    # ---- device-generated initial states ----
    def init_setup(self, name: str):
        """One of the reference's setups (include/benchmarks.hpp; names as in
        setups.SETUPS) evaluated on the device at this context's grid
        (mhd_init_setup)."""
        self._check(self.lib.mhd_init_setup(self.h, _setup_id(name)))

    def init_fourier(self, modes: np.ndarray):
        """Fourier ensemble member generated on the device (mhd_init_fourier)."""
        t = np.ascontiguousarray(modes, dtype=np.float64).reshape(-1, 8)
        self._check(self.lib.mhd_init_fourier(self.h, t.shape[0], _ptr(t)))

    def init_column(self, amps, phases, sausage: bool):
        """Kink / sausage column generated on the device (mhd_init_column)."""
        a = np.ascontiguousarray(amps, dtype=np.float64)
        p = np.ascontiguousarray(phases, dtype=np.float64)
        if a.shape != p.shape or a.ndim != 1:
            raise ValueError("amps and phases must be 1-D arrays of one length")
        self._check(self.lib.mhd_init_column(self.h, int(sausage), a.size, _ptr(a), _ptr(p)))

    # ---- files ----
    def write_snapshot(self, path: str, snapshot_rate: int, time_step: float):
        """write_data_to_binary_file (diagnostics.hpp:90-131) of the current state."""
        self._check(self.lib.mhd_write_snapshot(self.h, os.fsencode(path), snapshot_rate, time_step))

    def write_restart(self, path: str, t: float):
        self._check(self.lib.mhd_write_restart(self.h, os.fsencode(path), t))

    def read_restart(self, path: str) -> float:
        t = C.c_double()
        self._check(self.lib.mhd_read_restart(self.h, os.fsencode(path), C.byref(t)))
        return t.value

    def explicit_rhs_vjp(self, lam, fused: bool = True):
        """L'(U)^T lam for the explicit right-hand side at the current state;
        lam and the result are six padded arrays (rho, mx, my, E, bxf, byf).
        fused=False runs the four unfused kernels (the fused kernel's reference)."""
        keep, lp = self._field_ptrs(lam)
        outs = [self._padded() for _ in range(6)]
        op = (_dp * 6)(*[_ptr(o) for o in outs])
        fn = self.lib.mhd_explicit_rhs_vjp if fused else self.lib.mhd_explicit_rhs_vjp_unfused
        self._check(fn(self.h, lp, op))
        return outs

    def adjoint_forward(self, dt: float, nsteps: int, checkpoint_every: int = 10):
        """nsteps IMEX steps of constant dt from the current state, keeping
        checkpoints for adjoint_backward."""
        self._check(self.lib.mhd_adjoint_forward(self.h, dt, nsteps, checkpoint_every))

    def adjoint_backward(self, lam_final: dict):
        """Adjoint of the last adjoint_forward: given dJ/dU(final) (dict of six
        padded arrays), returns (dJ/dU(initial) as such a dict, dJ/d eta)."""
        names = ("rho", "mx", "my", "E", "bxf", "byf")
        keep, lp = self._field_ptrs([lam_final[k] for k in names])
        outs = {k: self._padded() for k in names}
        op = (_dp * 6)(*[_ptr(outs[k]) for k in names])
        eb = C.c_double()
        self._check(self.lib.mhd_adjoint_backward(self.h, lp, op, C.byref(eb)))
        return outs, eb.value

    def adjoint_use_graph(self, on: bool = True):
        """Replay the backward pass as a captured HIP graph (default) or
        launch its kernels directly (same kernels and order, same bits)."""
        self._check(self.lib.mhd_adjoint_use_graph(self.h, int(bool(on))))

    def adjoint_record_branches(self, on: bool = True):
        """Record, in the next adjoint_backward, every branch choice the
        adjoint linearises at (minmod, HLL case and wave-speed min / max,
        corner E_z upwinding, floors), for adjoint_forward_frozen."""
        self._check(self.lib.mhd_adjoint_record_branches(self.h, int(bool(on))))

    def adjoint_forward_frozen(self, nsteps: int, eta: float | None = None):
        """nsteps steps of the recorded dt from the current state with the
        recorded branch choices and exact implicit solves: the smooth map whose
        derivative the adjoint computes (eta: resistivity, default the
        context's)."""
        self._check(self.lib.mhd_adjoint_forward_frozen(self.h, nsteps, 0.0 if eta is None else float(eta)))

    # ---- instrumentation ----
    def profile(self, kernels=None):
        """Bracket launches of the named kernels (all if True, none if
        False/None) with HIP events; counters restart from zero."""
        if kernels is True:
            mask = (1 << len(_lib.KERNEL_IDS)) - 1
        elif not kernels:
            mask = 0
        else:
            mask = 0
            for k in ([kernels] if isinstance(kernels, str) else kernels):
                mask |= 1 << _lib.KERNEL_IDS[k]
        self._check(self.lib.mhd_profile_enable(self.h, mask))

    def profile_get(self, kernel: str):
        n, ms = C.c_longlong(), C.c_double()
        self._check(self.lib.mhd_profile_get(self.h, _lib.KERNEL_IDS[kernel], C.byref(n), C.byref(ms)))
        return n.value, ms.value

    def synchronize(self):
        self._check(self.lib.mhd_synchronize(self.h))

    def calibration_copy(self, nbytes: int):
        self._check(self.lib.mhd_calibration_copy(self.h, nbytes))

    def calibration_mix(self, n_in: int, n_out: int, n: int, reps: int = 5) -> float:
        """Mean ms per launch of a plain kernel reading n_in and writing
        n_out arrays of n doubles (mhd_calibration_mix)."""
        ms = C.c_double()
        self._check(self.lib.mhd_calibration_mix(self.h, n_in, n_out, n, reps, C.byref(ms)))
        return ms.value


class Batch:
    """Batched ensemble of independent periodic runs (mhd_batch_* of the C
    ABI): one launch per kernel advances every run; each run follows the
    reference driver loop with its own CFL time step, end time and operator."""

    FIELDS = ("rho", "mx", "my", "E", "bxf", "byf")

    def __init__(self, nx, ny, dx, dy, eta, *, cfl: float = 0.45, cg_tol: float | None = None,
                 cg_max_iter: int = 0, cg_fail_is_error: bool = False, march_rows: int = 0,
                 bc: tuple = ("periodic", "periodic")):
        """bc: boundary of the x and y axes of every run, "periodic" or
        "outflow" (the batched adjoint takes periodic runs only)."""
        self.lib = _lib.load()
        self.nx = np.ascontiguousarray(nx, dtype=np.int32)
        self.ny = np.ascontiguousarray(ny, dtype=np.int32)
        self.dx = np.ascontiguousarray(dx, dtype=np.float64)
        self.dy = np.ascontiguousarray(dy, dtype=np.float64)
        self.eta = np.ascontiguousarray(eta, dtype=np.float64)
        self.nruns = len(self.nx)
        for a in (self.ny, self.dx, self.dy, self.eta):
            if len(a) != self.nruns:
                raise ValueError("run arrays differ in length")
        o = _lib.Opts()
        _lib.check(self.lib.mhd_default_opts(C.byref(o)))
        o.cfl = cfl
        if cg_tol is not None:
            o.cg_tol = cg_tol
        o.cg_max_iter = cg_max_iter
        o.cg_fail_is_error = int(cg_fail_is_error)
        o.march_rows = march_rows
        try:
            o.bc_x, o.bc_y = (_lib.BC_NAMES[b] for b in bc)
        except (KeyError, ValueError, TypeError):
            raise ValueError(f"bc must be a pair of 'periodic' / 'outflow', got {bc!r}") from None
        self.bc = tuple(bc)
        h = C.c_void_p()
        ip = C.POINTER(C.c_int)
        _lib.check_batch(self.lib.mhd_batch_create(self.nruns, self.nx.ctypes.data_as(ip), self.ny.ctypes.data_as(ip),
                                                   _ptr(self.dx), _ptr(self.dy), _ptr(self.eta), C.byref(o),
                                                   C.byref(h)), None)
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.mhd_batch_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        _lib.check_batch(rc, self.h)

    def set_schedule(self, t_end=None, max_steps=None, dt_fixed=None):
        def arr(v, dtype):
            if v is None:
                return None
            return np.ascontiguousarray(np.broadcast_to(np.asarray(v, dtype=dtype), (self.nruns,)))
        te, ms, df = arr(t_end, np.float64), arr(max_steps, np.int32), arr(dt_fixed, np.float64)
        self._keep = (te, ms, df)
        self._check(self.lib.mhd_batch_set_schedule(
            self.h, _ptr(te), ms.ctypes.data_as(C.POINTER(C.c_int)) if ms is not None else None, _ptr(df)))

    def upload(self, run: int, state: dict):
        shape = (int(self.ny[run]) + 4, int(self.nx[run]) + 4)
        arrs = [np.ascontiguousarray(state[k], dtype=np.float64) for k in self.FIELDS]
        for a in arrs:
            if a.shape != shape:
                raise ValueError(f"run {run}: expected padded shape {shape}, got {a.shape}")
        self._check(self.lib.mhd_batch_upload(self.h, run, *[_ptr(a) for a in arrs]))

    def init_setup(self, run: int, name: str):
        """One of the reference's setups generated on the device at the run's
        grid (mhd_batch_init_setup)."""
        self._check(self.lib.mhd_batch_init_setup(self.h, run, _setup_id(name)))

    def init_fourier(self, run: int, modes: np.ndarray):
        """Device-generated Fourier ensemble member (mhd_batch_init_fourier);
        modes: the (M, 8) table of setups.fourier_member_modes."""
        t = np.ascontiguousarray(modes, dtype=np.float64).reshape(-1, 8)
        self._check(self.lib.mhd_batch_init_fourier(self.h, run, t.shape[0], _ptr(t)))

    def init_column(self, run: int, amps, phases, sausage: bool):
        """Device-generated kink / sausage column (mhd_batch_init_column)."""
        a = np.ascontiguousarray(amps, dtype=np.float64)
        p = np.ascontiguousarray(phases, dtype=np.float64)
        if a.shape != p.shape or a.ndim != 1:
            raise ValueError("amps and phases must be 1-D arrays of one length")
        self._check(self.lib.mhd_batch_init_column(self.h, run, int(sausage), a.size, _ptr(a), _ptr(p)))

    def download(self, run: int) -> dict:
        shape = (int(self.ny[run]) + 4, int(self.nx[run]) + 4)
        out = {k: np.zeros(shape) for k in self.FIELDS}
        self._check(self.lib.mhd_batch_download(self.h, run, *[_ptr(out[k]) for k in self.FIELDS]))
        return out

    def step(self) -> int:
        act = C.c_int()
        self._check(self.lib.mhd_batch_step(self.h, C.byref(act)))
        return act.value

    def run(self, max_batch_steps: int = -1) -> int:
        n = C.c_int()
        self._check(self.lib.mhd_batch_run(self.h, max_batch_steps, C.byref(n)))
        return n.value

    def autotune(self, trial_steps: int = 2):
        """Pick the fastest launch shape for the uploaded ensemble; returns
        (stage rows, tile rows).  The state is left as it was."""
        sr, tr = C.c_int(), C.c_int()
        self._check(self.lib.mhd_batch_autotune(self.h, trial_steps, C.byref(sr), C.byref(tr)))
        return sr.value, tr.value

    def tune(self, group_steps: int = 16, spec_margin: int = 0, graph: int = 0, fused_cfl: bool = True):
        """Launch mode of run(): steps per device-resident group (0: one
        host-checked step at a time), extra launched CG iterations per solve,
        graph replay (0 / False: direct launches, 1 / True: captured pairs of
        steps, 2: each group as one captured graph), CFL partials fused into
        the final recombination.  Results are identical in every mode."""
        self._check(self.lib.mhd_batch_tune(self.h, group_steps, spec_margin, int(graph), int(bool(fused_cfl))))

    def tuning(self) -> dict:
        """The current launch shape and mode."""
        v = [C.c_int() for _ in range(5)]
        self._check(self.lib.mhd_batch_get_tuning(self.h, *[C.byref(x) for x in v]))
        return dict(zip(("stage_rows", "tile_rows", "group_steps", "graph", "fused_cfl"), (x.value for x in v)))

    def run_info(self, run: int):
        steps, t, dt = C.c_int(), C.c_double(), C.c_double()
        its = (C.c_int * 2)()
        self._check(self.lib.mhd_batch_run_info(self.h, run, C.byref(steps), C.byref(t), C.byref(dt), its))
        return steps.value, t.value, dt.value, (its[0], its[1])

    def synchronize(self):
        self._check(self.lib.mhd_batch_synchronize(self.h))

    # ---- adjoint ----
    def adjoint_forward(self, nsteps: int, checkpoint_every: int):
        """Record nsteps steps of every run (each with its fixed dt_fixed and
        eta > 0) for the backward pass; checkpoints every checkpoint_every
        steps.  The batch then holds the final states."""
        self._check(self.lib.mhd_batch_adjoint_forward(self.h, nsteps, checkpoint_every))

    def adjoint_set_lambda(self, run: int, lam_final: dict):
        """dJ/dU(final) of one run (dict of six padded arrays)."""
        shape = (int(self.ny[run]) + 4, int(self.nx[run]) + 4)
        arrs = [np.ascontiguousarray(lam_final[k], dtype=np.float64) for k in self.FIELDS]
        for a in arrs:
            if a.shape != shape:
                raise ValueError(f"run {run}: expected padded shape {shape}, got {a.shape}")
        lp = (_dp * 6)(*[_ptr(a) for a in arrs])
        self._check(self.lib.mhd_batch_adjoint_set_lambda(self.h, run, lp))

    def adjoint_backward(self):
        """Backward pass of every run (consumes the final-state adjoints)."""
        self._check(self.lib.mhd_batch_adjoint_backward(self.h))

    def adjoint_result(self, run: int):
        """(dJ/dU(initial) as a dict of padded arrays, dJ/d eta) of one run."""
        shape = (int(self.ny[run]) + 4, int(self.nx[run]) + 4)
        outs = {k: np.zeros(shape) for k in self.FIELDS}
        op = (_dp * 6)(*[_ptr(outs[k]) for k in self.FIELDS])
        eb = C.c_double()
        self._check(self.lib.mhd_batch_adjoint_result(self.h, run, op, C.byref(eb)))
        return outs, eb.value


def operator_alpha(dt: float, eta: float) -> float:
    """alpha of generate_A_matrix as the driver forms it:
    (dt * resistivity) * A_im[4] (src/integrator_main.cpp:58-59)."""
    return (dt * eta) * GAMMA_ARS
