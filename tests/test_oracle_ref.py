"""Pins the oracle against the reference's own code, bit for bit.

oracle/_ref/libref_partial.so is compiled (oracle/Makefile) from the
reference's include/Grid.hpp, MUSCL.hpp, HLL.hpp and diagnostics.hpp -- the
parts of the reference that build without Eigen -- with the same strict flags
as the oracle (no fast-math, no FMA contraction).  With identical arithmetic
the two must agree exactly, so any dropped term, swapped index or wrong sign
in the restatement fails here.
"""
import ctypes as C
import os

import numpy as np
import pytest

from oracle import pyoracle as po

REF_SO = os.path.join(os.path.dirname(po.__file__), "_ref", "libref_partial.so")
pytestmark = pytest.mark.skipif(not os.path.exists(REF_SO), reason="oracle/_ref not built (reference tree absent)")

_dp = C.POINTER(C.c_double)


@pytest.fixture(scope="module")
def ref():
    return po.RefPartial(REF_SO)


@pytest.fixture(scope="module")
def lib():
    return po.load("strict")


def _arr6(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


def _random_state(rng, n, supersonic=None):
    """n random primitive states (rho, vx, vy, p, bx, by)."""
    s = np.empty((n, 6))
    s[:, 0] = rng.uniform(0.05, 3.0, n)
    s[:, 1] = rng.normal(0, 1.5, n)
    s[:, 2] = rng.normal(0, 1.5, n)
    s[:, 3] = rng.uniform(0.01, 3.0, n)
    s[:, 4] = rng.normal(0, 1.0, n)
    s[:, 5] = rng.normal(0, 1.0, n)
    if supersonic is not None:
        s[:, 1] = supersonic * rng.uniform(8, 12, n)
        s[:, 2] = supersonic * rng.uniform(8, 12, n)
    return s


def test_minmod(ref, lib):
    rng = np.random.default_rng(0)
    vals = list(rng.normal(size=200)) + [0.0, -0.0, 1.0, -1.0, 2.0, -2.0, 1e-300, -1e-300]
    for a in vals:
        for b in vals[::7]:
            assert lib.og_minmod(a, b) == ref.lib.ref_minmod(a, b)
    # ties return b, opposite signs and zeros return 0
    assert lib.og_minmod(1.0, 1.0) == 1.0
    assert lib.og_minmod(-2.0, 3.0) == 0.0
    assert lib.og_minmod(0.0, 3.0) == 0.0
    assert lib.og_minmod(0.5, 3.0) == 0.5
    assert lib.og_minmod(-3.0, -0.5) == -0.5


@pytest.mark.parametrize("branch", [None, 1.0, -1.0])
def test_hll_and_fast_speed(ref, lib, branch):
    rng = np.random.default_rng(1 if branch is None else (2 if branch > 0 else 3))
    L = _random_state(rng, 400, branch)
    R = _random_state(rng, 400, branch)
    R[:, 4] = L[:, 4]          # x faces: common normal field
    out_o = np.zeros(6); out_r = np.zeros(6)
    po_o = out_o.ctypes.data_as(_dp); po_r = out_r.ctypes.data_as(_dp)
    for k in range(L.shape[0]):
        l, lp = _arr6(L[k]); r, rp = _arr6(R[k])
        assert lib.og_fast_speed(lp, l[4]) == ref.lib.ref_fast_speed(lp, l[4])
        lib.og_hll_x(lp, rp, po_o); ref.lib.ref_hll_x(lp, rp, po_r)
        assert np.array_equal(out_o, out_r), (k, out_o, out_r)
        r2 = r.copy(); r2[5] = l[5]; r2[4] = R[k, 4]
        r2c, r2p = _arr6(r2)
        lib.og_hll_y(lp, r2p, po_o); ref.lib.ref_hll_y(lp, r2p, po_r)
        assert np.array_equal(out_o, out_r), (k, out_o, out_r)


def _random_padded(rng, nx, ny):
    shape = (ny + 4, nx + 4)
    f = {
        "rho": rng.uniform(0.1, 2.0, shape), "mx": rng.normal(0, 0.5, shape), "my": rng.normal(0, 0.5, shape),
        "E": rng.uniform(2.0, 6.0, shape), "bxf": rng.normal(0, 1.0, shape), "byf": rng.normal(0, 1.0, shape),
    }
    # floor-triggering cells: tiny density, negative pressure
    f["rho"][5, 6] = 1e-12
    f["E"][7, 3] = -1.0
    return f


def _all12(f, nx, ny):
    shape = (ny + 4, nx + 4)
    arrs = [np.ascontiguousarray(f.get(k, np.zeros(shape)), dtype=np.float64).copy() for k in po.FIELDS]
    return arrs


@pytest.mark.parametrize("nx,ny", [(8, 8), (13, 6), (32, 20)])
def test_ghost_fill_and_primitives(ref, nx, ny):
    rng = np.random.default_rng(nx * 100 + ny)
    f = _random_padded(rng, nx, ny)
    arrs_r = _all12(f, nx, ny)
    ref.lib.ref_fill_and_primitives(nx, ny, ref.ptrs(arrs_r))
    s = po.Sim(nx, ny, 1.0, 1.0, 0.0)
    for k, a in zip(po.FIELDS, _all12(f, nx, ny)):
        s.field(k)[...] = a
    s.fill_ghost()
    s.primitives()
    for k, a in zip(po.FIELDS, arrs_r):
        if k == "Ez":
            continue
        assert np.array_equal(s.field(k), a), k


@pytest.mark.parametrize("nx,ny", [(8, 8), (11, 17)])
def test_muscl_all(ref, nx, ny):
    rng = np.random.default_rng(7 + nx)
    f = _random_padded(rng, nx, ny)
    arrs = _all12(f, nx, ny)
    ref.lib.ref_fill_and_primitives(nx, ny, ref.ptrs(arrs))
    rv_r = [np.zeros((ny + 4, nx + 4)) for _ in range(20)]
    ref.lib.ref_muscl_all(nx, ny, ref.ptrs(arrs), ref.ptrs(rv_r))
    s = po.Sim(nx, ny, 1.0, 1.0, 0.0)
    for k, a in zip(po.FIELDS, arrs):
        s.field(k)[...] = a
    s.lib.og_muscl_all(s.g, s.w)
    for q in range(20):
        got = np.ctypeslib.as_array(s.lib.og_ws_rv(s.w, q), shape=(ny + 4, nx + 4))
        assert np.array_equal(got, rv_r[q]), q


@pytest.mark.parametrize("eta,cfl", [(0.0, 0.45), (1e-3, 0.45), (5.0, 0.45)])
def test_cfl_divb(ref, eta, cfl):
    nx, ny = 24, 18
    rng = np.random.default_rng(11)
    f = _random_padded(rng, nx, ny)
    arrs = _all12(f, nx, ny)
    ref.lib.ref_fill_and_primitives(nx, ny, ref.ptrs(arrs))
    dx, dy = 0.3, 0.25
    s = po.Sim(nx, ny, dx, dy, eta)
    for k, a in zip(po.FIELDS, arrs):
        s.field(k)[...] = a
    # the reference hard-codes an effective Courant number of 0.45
    assert s.cfl(cfl) == ref.lib.ref_cfl(nx, ny, dx, dy, eta, ref.ptrs(arrs))
    rms, mx = C.c_double(), C.c_double()
    ref.lib.ref_divb(nx, ny, dx, dy, ref.ptrs(arrs), C.byref(rms), C.byref(mx))
    orms, omx = s.divb()
    assert (orms, omx) == (rms.value, mx.value)
