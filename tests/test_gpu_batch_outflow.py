"""GPU parity of batched ensembles with outflow (zero-gradient) boundaries.

A batch takes the boundaries of its options for every run.  Each run of an
outflow batch is checked against the CPU oracle on the same inputs and
boundaries (relative L1 per field <= 1e-10, time steps to round-off, CG
iterations within +-1), downloads carry the outflow ghosts, the
device-resident launch modes equal the host-checked step path bit for bit,
and the batched adjoint refuses an outflow batch.
"""
import numpy as np
import pytest

from helpers import CONSERVED, ic_arrays, interior, rel_l1
from magnetohydrodynamics_simulator_amd import Batch
from magnetohydrodynamics_simulator_amd._lib import MhdError
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

OUT = ("outflow", "outflow")
TOL_RUN = 1e-10

# (setup, nx, ny, eta): uneven sizes, an ideal run among resistive ones
RUNS = [
    ("orszag_tang", 64, 48, 1e-2),
    ("orszag_tang", 40, 72, 3e-3),
    ("kink", 70, 44, 1e-3),
    ("orszag_tang", 36, 36, 0.0),
]


def _say(msg):
    print(f"\n[case] {msg}", flush=True)


def _states():
    out = []
    for k, (name, nx, ny, eta) in enumerate(RUNS):
        ic, dx, dy = ic_arrays(name, nx, ny, seed=30 + k, amp=1e-3)
        out.append((ic, nx, ny, dx, dy, eta))
    return out


def _batch(states, bc, tol=1e-10, cfl=0.4):
    b = Batch([s[1] for s in states], [s[2] for s in states], [s[3] for s in states], [s[4] for s in states],
              [s[5] for s in states], cfl=cfl, cg_tol=tol, bc=bc)
    for r, s in enumerate(states):
        b.upload(r, s[0])
    return b


@pytest.mark.parametrize("bc", [OUT, ("outflow", "periodic"), ("periodic", "outflow")],
                         ids=["open-xy", "open-x", "open-y"])
def test_outflow_batch_matches_oracle_per_run(bc):
    _say(f"outflow batch of {len(RUNS)} uneven runs bc={bc}, 6 steps, vs oracle")
    tol, cfl, steps = 1e-10, 0.4, 6
    states = _states()
    with _batch(states, bc, tol, cfl) as b:
        b.set_schedule(max_steps=steps)
        assert b.run() == steps
        for r, (ic, nx, ny, dx, dy, eta) in enumerate(states):
            got = b.download(r)
            nsteps, t, dt, its = b.run_info(r)
            sim = po.Sim(nx, ny, dx, dy, eta, bc=bc)
            sim.set_conserved({k: ic[k] for k in CONSERVED})
            sim.set_cg(tol, -1)
            n, dts, oits, ot = sim.run(1e9, cfl=cfl, max_steps=steps)
            assert nsteps == n == steps
            assert abs(t - ot) <= 1e-13 * ot and abs(dt - dts[-1]) <= 1e-12 * dts[-1], (r, dt, dts[-1])
            assert abs(its[0] - oits[-1][0]) <= 1 and abs(its[1] - oits[-1][1]) <= 1, (r, its, oits[-1])
            for k in CONSERVED:
                e = rel_l1(interior(got[k]), sim.interior(k))
                assert e <= TOL_RUN, (r, k, e)
                # ghosts: on an open axis the nearest inside value, else periodic
                f = got[k]
                if bc[0] == "outflow":
                    assert np.array_equal(f[2:-2, 0], f[2:-2, 2]) and np.array_equal(f[2:-2, -1], f[2:-2, -3])
                else:
                    assert np.array_equal(f[2:-2, 0], f[2:-2, -4]) and np.array_equal(f[2:-2, -1], f[2:-2, 3])
                if bc[1] == "outflow":
                    assert np.array_equal(f[0], f[2]) and np.array_equal(f[-1], f[-3])
                else:
                    assert np.array_equal(f[0], f[-4]) and np.array_equal(f[-1], f[3])
            sim.close()


MODES = [
    dict(group_steps=16),
    dict(group_steps=5, graph=True, fused_cfl=False),
    dict(group_steps=7, graph=2),
]


def test_outflow_batch_launch_modes_match_step_path():
    """Device-resident groups (CG budgets, halt and resume, graph replay,
    fused or separate CFL partials) of an outflow batch against
    mhd_batch_step in a loop: identical fields, steps, times and CG counts."""
    tol, cfl = 1e-10, 0.4
    states = _states()
    t_end = [0.02, 0.03, 0.015, 1e9]

    def run(mode):
        with _batch(states, OUT, tol, cfl) as b:
            b.tune(**mode)
            b.set_schedule(t_end=t_end)
            done = [b.run(n) for n in (8, 5)]
            return done, [b.download(r) for r in range(len(states))], [b.run_info(r) for r in range(len(states))]

    _say("outflow batch launch modes: host-checked steps (reference)")
    ref_done, ref, ref_info = run(dict(group_steps=0))
    for mode in MODES:
        _say(f"outflow batch launch mode {mode} vs host-checked steps")
        done, out, info = run(mode)
        assert done == ref_done, (done, ref_done)
        for r in range(len(states)):
            assert info[r][:3] == ref_info[r][:3] and list(info[r][3]) == list(ref_info[r][3]), (r, info[r])
            for k in CONSERVED:
                assert np.array_equal(out[r][k], ref[r][k]), (mode, r, k)


def test_outflow_batch_adjoint_refused():
    _say("batched adjoint of an outflow batch is refused")
    states = _states()[:2]
    with _batch(states, OUT) as b:
        b.set_schedule(max_steps=2, dt_fixed=[1e-3, 1e-3])
        with pytest.raises(MhdError, match="periodic"):
            b.adjoint_forward(2, 1)


def test_outflow_batch_rejects_unknown_boundary():
    with pytest.raises(ValueError):
        Batch([32], [32], [0.1], [0.1], [0.0], bc=("outflow", "reflecting"))
