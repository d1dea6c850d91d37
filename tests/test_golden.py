"""Golden fixtures (tests/golden/, written by tests/golden/gen_golden.py).

CPU: the strict oracle reproduces them bit for bit.
GPU: the HIP path, started from the stored inputs, matches them to the
single-run tolerance (relative L1 <= 1e-12 per field after 1 and 5 steps) with
the same CG iteration counts within +-1 and the same dt sequence.
"""
import os

import numpy as np
import pytest

from helpers import CONSERVED, interior

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = [(0.0, 1), (0.0, 5), (1e-2, 1), (1e-2, 5)]


def load(eta, steps):
    return dict(np.load(os.path.join(HERE, f"ot32_eta{eta:g}_steps{steps}.npz"), allow_pickle=False))


def padded_inputs(g, n=32):
    out = {}
    for k in CONSERVED:
        a = np.zeros((n + 4, n + 4))
        a[2:-2, 2:-2] = g[f"in_{k}"]
        out[k] = a
    return out


@pytest.mark.parametrize("eta,steps", CASES)
def test_oracle_reproduces_golden(eta, steps):
    import sys
    sys.path.insert(0, HERE)
    from gen_golden import run_case

    g = load(eta, steps)
    d = run_case(eta, steps)
    for k in CONSERVED:
        assert np.array_equal(d[k], g[k]), k
    assert np.array_equal(d["cg_iters"], g["cg_iters"])
    assert np.array_equal(d["dts"], g["dts"])


@pytest.mark.gpu
@pytest.mark.parametrize("eta,steps", CASES)
def test_gpu_matches_golden(eta, steps):
    from magnetohydrodynamics_simulator_amd import Context, setups

    print(f"\n[case] golden OT 32^2 eta={eta} steps={steps}", flush=True)
    g = load(eta, steps)
    n = 32
    dx = setups.grid_spacing(n)
    ic = padded_inputs(g)
    with Context(n, n, dx, dx, eta, cfl=0.45) as ctx:
        ctx.upload(*[ic[k] for k in CONSERVED])
        nsteps, dts, its, _ = ctx.run(1e9, max_steps=steps)
        out = ctx.download()
    assert nsteps == steps
    np.testing.assert_allclose(dts, g["dts"], rtol=1e-13, atol=0)
    assert np.abs(its - g["cg_iters"]).max() <= 1
    for k in CONSERVED:
        a, b = interior(out[k]), g[k]
        assert np.abs(a - b).sum() / np.abs(b).sum() <= 1e-12, k
