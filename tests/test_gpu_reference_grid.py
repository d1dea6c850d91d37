"""The C++ adapter (include/mhd_mi355x.hpp) driven from a call site built on
the reference's own Grid (include/Grid.hpp), the way the reference's main
drives its step (src/integrator_main.cpp:34-39, 52-112).

tests/cpp/reference_grid_call_site is built by tests/cpp/Makefile where the
reference tree is present and ships prebuilt with the tree; the test skips
where it is absent.  Each adapter call uploads the Grid, runs the library's
entry point and downloads into the Grid, so the binary's results must equal,
bit for bit, the same calls made on the same state through the Python
Context (which the oracle parity tests check): explicit_rhs,
CFL_condition_checker, generate_A_matrix + IMEX, resistive_B_term,
calculate_current_div_B and integrate (mhd_run_resident; with a snapshot frequency, mhd_run_resident_files,
whose files must equal those a callback writes with mhd_write_snapshot).
"""
import os
import subprocess

import numpy as np
import pytest

from helpers import CONSERVED
from magnetohydrodynamics_simulator_amd import Context, operator_alpha, setups

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "_bin", "reference_grid_call_site")


@pytest.mark.parametrize("n,eta,t_end,every", [(64, 1e-2, 0.05, 0), (96, 0.0, 0.03, 0), (64, 1e-2, 0.05, 4)])
def test_reference_grid_call_site_matches_context(tmp_path, n, eta, t_end, every):
    if not os.path.exists(BIN):
        pytest.skip("tests/cpp/_bin/reference_grid_call_site not built (needs the reference tree)")
    print(f"\n[case] reference Grid call site, Orszag-Tang {n}^2, eta {eta}, t_end {t_end}", flush=True)
    dx = setups.grid_spacing(n)
    ic = setups.orszag_tang(n)
    P = (n + 4) * (n + 4)
    hdr = np.array([n, n, dx, dx, eta, t_end], dtype=np.float64)
    inp = np.concatenate([hdr] + [np.ascontiguousarray(ic[k], dtype=np.float64).ravel() for k in CONSERVED])
    inp.tofile(tmp_path / "in.bin")
    args = [BIN, str(tmp_path / "in.bin"), str(tmp_path / "out.bin")] + ([str(every)] if every else [])
    (tmp_path / "output").mkdir()
    r = subprocess.run(args, capture_output=True, text=True, timeout=120, cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    out = np.fromfile(tmp_path / "out.bin", dtype=np.float64)
    assert out.size == 4 + 20 * P
    dt_b, steps_b, rms_b, max_b = out[:4]
    arrs = out[4:].reshape(20, n + 4, n + 4)
    rhs_b, res_b, step_b, end_b = arrs[:6], arrs[6:8], arrs[8:14], arrs[14:]

    def up(c, st):
        c.upload(*[st[k] for k in CONSERVED])

    with Context(n, n, dx, dx, eta, cfl=0.4, cg_tol=1e-10) as c:
        up(c, ic)
        rhs = c.explicit_rhs()
        dt = c.CFL_condition_checker()
        c.generate_A_matrix(operator_alpha(dt, eta))
        c.IMEX(dt)
        stepped = c.download()
        up(c, stepped)
        bx_rate, by_rate, _ = c.resistive_B_term(True)
        rms, mx = c.calculate_current_div_B()
        up(c, stepped)
        ref_dir = tmp_path / "ref"
        ref_dir.mkdir()

        def on_snap(ctx, frame, t, dt):
            ctx.write_snapshot(os.fsdecode(ref_dir / f"snapshot_{frame:06d}.dat"), every, dt)

        steps, _, _, _ = c.run_resident(t_end, snapshot_every=every, on_snapshot=on_snap if every else None)
        final = c.download()
    print(f"[case] dt {dt_b!r} / {dt!r}, steps {int(steps_b)} / {steps}, div B max {max_b:.3e}", flush=True)
    assert dt_b == dt and int(steps_b) == steps and rms_b == rms and max_b == mx
    for q in range(6):
        assert np.array_equal(rhs_b[q], rhs[q]), ("explicit_rhs", CONSERVED[q])
        assert np.array_equal(step_b[q], stepped[CONSERVED[q]]), ("IMEX", CONSERVED[q])
        assert np.array_equal(end_b[q], final[CONSERVED[q]]), ("integrate", CONSERVED[q])
    assert np.array_equal(res_b[0], bx_rate) and np.array_equal(res_b[1], by_rate)
    assert steps > 1 and max_b < 1e-11
    # integrate's snapshot files (written while the loop runs on) = the callback's
    names = sorted(p.name for p in (tmp_path / "output").iterdir())
    assert names == sorted(p.name for p in (tmp_path / "ref").iterdir())
    assert len(names) == (1 + steps // every if every else 0)
    for name in names:
        assert (tmp_path / "output" / name).read_bytes() == (tmp_path / "ref" / name).read_bytes(), name
