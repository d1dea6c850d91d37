"""GPU parity of outflow (zero-gradient) boundaries against the oracle.

The reference is periodic only; the outflow rule (include/mhd_mi355x.h,
og_grid_set_bc of the oracle) is pinned on the CPU by
test_oracle_outflow.py.  Here the library must reproduce the oracle on the
same inputs, to the bars of the periodic parity tests: the explicit right-hand
side and the resistive terms to round-off, whole runs to 1e-10 relative L1
with the same CFL steps and CG counts within one.  Cases put structure at
every boundary (an Orszag-Tang state, a perturbed one, a shock tube whose
waves leave the box) with both axes open, one axis open either way, and
grids that are not multiples of the tile or strip widths.  The resident loop
must equal mhd_run bitwise here as well (its final recombination forms the
CFL partials at the open top boundary), downloads carry the outflow ghosts,
and the periodic-only paths refuse outflow contexts.
"""
import ctypes as C

import numpy as np
import pytest

from helpers import CONSERVED, ic_arrays, interior, rel_l1
from magnetohydrodynamics_simulator_amd import Context, _lib, setups
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu
OUT = ("outflow", "outflow")
TOL_RUN = 1e-10


def _say(msg):
    print(f"\n[case] {msg}", flush=True)


def _pair(ic, nx, ny, dx, dy, eta, bc, tol=1e-12, cfl=0.4):
    sim = po.Sim(nx, ny, dx, dy, eta, bc=bc)
    sim.set_conserved({k: ic[k] for k in CONSERVED})
    sim.set_cg(tol, -1)
    ctx = Context(nx, ny, dx, dy, eta, cfl=cfl, cg_tol=tol, bc=bc)
    ctx.upload(*[ic[k] for k in CONSERVED])
    return sim, ctx


@pytest.mark.parametrize("bc", [OUT, ("outflow", "periodic"), ("periodic", "outflow")],
                         ids=["open-xy", "open-x", "open-y"])
def test_explicit_rhs_and_resistive_terms(bc):
    nx, ny = 70, 52
    _say(f"explicit rhs / resistive terms {nx}x{ny} bc={bc}")
    ic, dx, dy = ic_arrays("orszag_tang", nx, ny, seed=8, amp=1e-2)
    sim, ctx = _pair(ic, nx, ny, dx, dy, 2e-2, bc, tol=1e-14)
    with ctx:
        got = ctx.explicit_rhs()
        ref = sim.explicit_rhs()
        for q, k in enumerate(CONSERVED):
            assert rel_l1(interior(got[q]), interior(ref[q])) <= 1e-13, k
        gx, gy, _ = ctx.resistive_B_term(False)
        rx, ry, _ = sim.resistive_term(False)
        assert rel_l1(interior(gx), interior(rx)) <= 1e-12
        assert rel_l1(interior(gy), interior(ry)) <= 1e-12
        alpha = 2e-3 * 2e-2 * 0.2928932188134524
        ctx.generate_A_matrix(alpha)
        sim.build_operator(alpha)
        gx, gy, git = ctx.resistive_B_term(True)
        rx, ry, rit = sim.resistive_term(True)
        assert abs(git - rit) <= 1
        assert rel_l1(interior(gx), interior(rx)) <= 1e-9
        assert rel_l1(interior(gy), interior(ry)) <= 1e-9
    sim.close()


@pytest.mark.parametrize("bc,nx,ny,eta,steps", [
    (OUT, 64, 64, 1e-2, 8),
    (("outflow", "periodic"), 80, 48, 3e-3, 8),
    (("periodic", "outflow"), 48, 80, 3e-3, 8),
    (OUT, 300, 40, 1e-3, 5),
], ids=["open-xy-64", "open-x-80x48", "open-y-48x80", "open-xy-300x40"])
def test_runs_match_oracle(bc, nx, ny, eta, steps):
    _say(f"OT run {nx}x{ny} eta={eta} bc={bc}, {steps} steps")
    ic, dx, dy = ic_arrays("orszag_tang", nx, ny, seed=2, amp=1e-3)
    sim, ctx = _pair(ic, nx, ny, dx, dy, eta, bc, tol=1e-12)
    nr, dts_r, it_r, _ = sim.run(10.0, cfl=0.4, max_steps=steps)
    with ctx:
        ng, dts_g, it_g, _ = ctx.run(10.0, max_steps=steps)
        got = ctx.download()
        rms, _ = ctx.calculate_current_div_B()
    assert ng == nr == steps
    assert np.allclose(dts_g, dts_r, rtol=1e-12, atol=0)
    assert np.abs(it_g - it_r).max() <= 1, (it_g, it_r)
    for k in CONSERVED:
        e = rel_l1(interior(got[k]), sim.interior(k))
        assert e <= TOL_RUN, (k, e)
    sr, _ = sim.divb()
    assert abs(rms - sr) <= 1e-9 * max(sr, 1e-12)
    sim.close()


@pytest.mark.parametrize("eta", [0.0, 1e-3])
def test_shock_tube_through_the_boundary(eta):
    nx, ny = 128, 8
    _say(f"shock tube {nx}x{ny} eta={eta}, outflow x, to t = 0.2")
    ic = setups.shock_tube(nx, ny)
    dx = 1.0 / nx
    bc = ("outflow", "periodic")
    sim, ctx = _pair(ic, nx, ny, dx, dx, eta, bc)
    nr, dts_r, _, tr = sim.run(0.2, cfl=0.4)
    with ctx:
        ng, dts_g, _, tg = ctx.run(0.2)
        got = ctx.download()
    assert ng == nr and tg == tr
    assert np.allclose(dts_g, dts_r, rtol=1e-12, atol=0)
    for k in CONSERVED:
        a = interior(got[k])
        assert np.all(a == a[0:1, :]), k          # one-dimensional, bit for bit
        assert rel_l1(a, sim.interior(k)) <= TOL_RUN, k
    rho = interior(got["rho"])[0]
    assert rho[-1] != 0.125                       # the fast wave has left through x = 1
    sim.close()


def test_resident_loop_equals_run_with_outflow():
    nx, ny, eta = 96, 70, 5e-3
    _say(f"resident vs mhd_run {nx}x{ny} eta={eta} open x and y")
    ic, dx, dy = ic_arrays("orszag_tang", nx, ny, seed=4, amp=1e-3)
    outs = []
    for resident in (False, True):
        with Context(nx, ny, dx, dy, eta, cfl=0.4, cg_tol=1e-10, bc=OUT) as c:
            c.upload(*[ic[k] for k in CONSERVED])
            res = (c.run_resident if resident else c.run)(10.0, max_steps=12)
            outs.append((res, c.download()))
    (n0, d0, i0, t0), f0 = outs[0]
    (n1, d1, i1, t1), f1 = outs[1]
    assert n0 == n1 and t0 == t1 and np.array_equal(d0, d1) and np.array_equal(i0, i1)
    for k in CONSERVED:
        assert np.array_equal(f0[k], f1[k]), k


def test_download_ghosts_and_periodic_only_paths():
    nx, ny = 20, 16
    _say("outflow ghosts of a download; the periodic-only adjoint refuses outflow; boundary codes checked")
    ic, dx, dy = ic_arrays("orszag_tang", nx, ny, seed=1, amp=1e-2)
    with Context(nx, ny, dx, dy, 1e-3, bc=OUT) as c:
        c.upload(*[ic[k] for k in CONSERVED])
        d = c.download()
        for k in CONSERVED:
            f = d[k]
            assert np.array_equal(f[2:-2, 0], f[2:-2, 2]) and np.array_equal(f[2:-2, 1], f[2:-2, 2])
            assert np.array_equal(f[2:-2, -1], f[2:-2, -3]) and np.array_equal(f[0], f[2])
            assert np.array_equal(f[-1], f[-3])
        with pytest.raises(RuntimeError, match="periodic"):
            c.adjoint_forward(1e-3, 2, 1)
    lib = _lib.load()
    o = _lib.Opts()
    lib.mhd_default_opts(C.byref(o))
    n = (C.c_int * 1)(32)
    h = (C.c_double * 1)(0.1)
    e = (C.c_double * 1)(1e-3)
    b = C.c_void_p()
    o.bc_x = 7                                   # neither periodic nor outflow
    assert lib.mhd_batch_create(1, n, n, h, h, e, C.byref(o), C.byref(b)) == _lib.MHD_EINVAL
    o.bc_x = _lib.BC_OUTFLOW                     # batches take outflow axes (tests/test_gpu_batch_outflow.py)
    assert lib.mhd_batch_create(1, n, n, h, h, e, C.byref(o), C.byref(b)) == _lib.MHD_OK
    assert lib.mhd_batch_destroy(b) == _lib.MHD_OK
    with pytest.raises(ValueError):
        Context(nx, ny, dx, dy, bc=("outflow", "reflecting"))
