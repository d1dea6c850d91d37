"""Seeded random sweep of the GPU driver loop against the oracle's.

Each case draws a grid shape (sides 5..160, aspect ratios up to 1:12), one
of the reference's setups (include/benchmarks.hpp) with a random
divergence-free perturbation, a resistivity (zero in a quarter of the
cases, else log-uniform in [1e-4, 0.3]), a Courant number in [0.1, 0.45],
a CG tolerance in [1e-12, 1e-8] and, in a third of the cases, outflow
boundaries on one or both axes; then runs three driver steps
(src/integrator_main.cpp:52-112) on the GPU (mhd_run) and in the oracle.
Bar: the same time steps (1e-12 relative), CG iteration counts within one,
relative L1 <= 1e-10 per conserved field, and on periodic grids |div B|max
at the oracle's round-off level (at most twice its value plus 1e-13: it
scales with |B| / h, 1e-15..1e-11 across these setups).  With outflow axes
the periodic setups are not divergence-free at the boundary, so div B is
not checked there.
"""
import numpy as np
import pytest

from helpers import CONSERVED, ic_arrays, interior, make_oracle, rel_l1
from magnetohydrodynamics_simulator_amd import Context

pytestmark = pytest.mark.gpu

SETUP_NAMES = ("orszag_tang", "decaying_sine", "alfven", "gaussian", "acoustic", "kink", "sausage")
CASES = 24


def _draw(seed):
    rng = np.random.default_rng(1000 + seed)
    nx = int(rng.integers(5, 161))
    ny = int(np.clip(round(nx * np.exp(rng.uniform(-2.5, 2.5))), 5, 160))
    name = SETUP_NAMES[seed % len(SETUP_NAMES)]
    eta = 0.0 if rng.random() < 0.25 else float(np.exp(rng.uniform(np.log(1e-4), np.log(0.3))))
    cfl = float(rng.uniform(0.1, 0.45))
    tol = float(np.exp(rng.uniform(np.log(1e-12), np.log(1e-8))))
    r = rng.random()
    bc = ("periodic", "periodic") if r < 2 / 3 else (("outflow", "periodic"), ("periodic", "outflow"),
                                                      ("outflow", "outflow"))[int(rng.integers(0, 3))]
    return nx, ny, name, eta, cfl, tol, bc


@pytest.mark.parametrize("seed", range(CASES))
def test_random_case_matches_oracle(seed):
    nx, ny, name, eta, cfl, tol, bc = _draw(seed)
    print(f"\n[case] seed {seed}: {name} {nx}x{ny} eta={eta:.3g} cfl={cfl:.3f} tol={tol:.2g} bc={bc}", flush=True)
    ic, dx, dy = ic_arrays(name, nx, ny, seed=seed, amp=1e-3)
    sim = make_oracle(ic, nx, ny, dx, dy, eta, tol=tol)
    if bc != ("periodic", "periodic"):
        sim.close()
        from oracle import pyoracle as po
        sim = po.Sim(nx, ny, dx, dy, eta, bc=bc)
        for k in CONSERVED:
            sim.field(k)[...] = ic[k]
        sim.set_cg(tol, -1)
    nr, dts_r, its_r, _ = sim.run(1e9, cfl=cfl, max_steps=3, snapshot_every=0)
    with Context(nx, ny, dx, dy, eta, cfl=cfl, cg_tol=tol, bc=bc) as c:
        c.upload(*[ic[k] for k in CONSERVED])
        ng, dts_g, its_g, _ = c.run(1e9, max_steps=3)
        got = c.download()
        _, divmax = c.calculate_current_div_B()
    assert ng == nr == 3
    assert np.allclose(dts_g, dts_r, rtol=1e-12, atol=0), (dts_g, dts_r)
    assert np.abs(its_g - its_r).max() <= 1, (its_g, its_r)
    for k in CONSERVED:
        e = rel_l1(interior(got[k]), sim.interior(k))
        assert e <= 1e-10, (k, e)
    if bc == ("periodic", "periodic"):
        _, div_ref = sim.divb()
        assert divmax <= 2 * div_ref + 1e-13, (divmax, div_ref)
    sim.close()
