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
The call site's --setup mode fills the Grid through the adapter's
test_selection (the reference's seven setups generated on the device).
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


def _run_context_sequence(tmp_path, nx, ny, dx, dy, eta, t_end, every, ic=None, setup=None):
    """The call site's sequence through the Python Context; returns its results."""
    def up(c, st):
        c.upload(*[st[k] for k in CONSERVED])

    with Context(nx, ny, dx, dy, eta, cfl=0.4, cg_tol=1e-10) as c:
        if setup is None:
            up(c, ic)
        else:
            c.init_setup(setup)
        initial = c.download()
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
    return dict(initial=initial, rhs=rhs, dt=dt, stepped=stepped, bx_rate=bx_rate, by_rate=by_rate, rms=rms,
                mx=mx, steps=steps, final=final)


def _check_call_site(tmp_path, args, nx, ny, every, ref, initial_from_device):
    (tmp_path / "output").mkdir()
    r = subprocess.run(args, capture_output=True, text=True, timeout=120, cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    P = (nx + 4) * (ny + 4)
    out = np.fromfile(tmp_path / "out.bin", dtype=np.float64)
    assert out.size == 4 + 26 * P
    dt_b, steps_b, rms_b, max_b = out[:4]
    arrs = out[4:].reshape(26, ny + 4, nx + 4)
    rhs_b, res_b, step_b, end_b, init_b = arrs[:6], arrs[6:8], arrs[8:14], arrs[14:20], arrs[20:]
    steps = ref["steps"]
    print(f"[case] dt {dt_b!r} / {ref['dt']!r}, steps {int(steps_b)} / {steps}, div B max {max_b:.3e}", flush=True)
    assert dt_b == ref["dt"] and int(steps_b) == steps and rms_b == ref["rms"] and max_b == ref["mx"]
    for q in range(6):
        k = CONSERVED[q]
        if initial_from_device:   # (an input file's ghosts are the caller's)
            assert np.array_equal(init_b[q], ref["initial"][k]), ("initial state", k)
        assert np.array_equal(rhs_b[q], ref["rhs"][q]), ("explicit_rhs", k)
        assert np.array_equal(step_b[q], ref["stepped"][k]), ("IMEX", k)
        assert np.array_equal(end_b[q], ref["final"][k]), ("integrate", k)
    assert np.array_equal(res_b[0], ref["bx_rate"]) and np.array_equal(res_b[1], ref["by_rate"])
    assert steps > 1 and max_b < 1e-11
    # integrate's snapshot files (written while the loop runs on) = the callback's
    names = sorted(p.name for p in (tmp_path / "output").iterdir())
    assert names == sorted(p.name for p in (tmp_path / "ref").iterdir())
    assert len(names) == (1 + steps // every if every else 0)
    for name in names:
        assert (tmp_path / "output" / name).read_bytes() == (tmp_path / "ref" / name).read_bytes(), name
    return init_b


def _need_binary():
    if not os.path.exists(BIN):
        pytest.skip("tests/cpp/_bin/reference_grid_call_site not built (needs the reference tree)")


@pytest.mark.parametrize("n,eta,t_end,every", [(64, 1e-2, 0.05, 0), (96, 0.0, 0.03, 0), (64, 1e-2, 0.05, 4)])
def test_reference_grid_call_site_matches_context(tmp_path, n, eta, t_end, every):
    _need_binary()
    print(f"\n[case] reference Grid call site, Orszag-Tang {n}^2, eta {eta}, t_end {t_end}", flush=True)
    dx = setups.grid_spacing(n)
    ic = setups.orszag_tang(n)
    hdr = np.array([n, n, dx, dx, eta, t_end], dtype=np.float64)
    inp = np.concatenate([hdr] + [np.ascontiguousarray(ic[k], dtype=np.float64).ravel() for k in CONSERVED])
    inp.tofile(tmp_path / "in.bin")
    args = [BIN, str(tmp_path / "in.bin"), str(tmp_path / "out.bin")] + ([str(every)] if every else [])
    ref = _run_context_sequence(tmp_path, n, n, dx, dx, eta, t_end, every, ic=ic)
    init_b = _check_call_site(tmp_path, args, n, n, every, ref, False)
    for q, k in enumerate(CONSERVED):
        assert np.array_equal(init_b[q], ic[k]), k


# test_selection's menu (benchmarks.hpp:424-430) -> setups.SETUPS names
MENU = {1: "orszag_tang", 2: "sausage", 3: "kink", 4: "decaying_sine", 5: "alfven", 6: "gaussian", 7: "acoustic"}


@pytest.mark.parametrize("choice", sorted(MENU))
def test_reference_grid_test_selection(tmp_path, choice):
    """main's set-up (integrator_main.cpp:34-50) on a 48 x 40 Grid with the
    setup picked by test_selection, generated on the device (mhd_init_setup):
    the call site's whole sequence equals the Context's from init_setup bit
    for bit, and the initial state equals the host restatement of the
    reference's setup (setups.py) to 1e-13 of each field's maximum."""
    _need_binary()
    nx, ny, eta, t_end, every = 48, 40, 5e-3, 0.12, 3
    name = MENU[choice]
    print(f"\n[case] reference Grid call site, test_selection {choice} ({name}) {nx}x{ny}", flush=True)
    dx, dy = setups.grid_spacing(nx), setups.grid_spacing(ny)
    args = [BIN, "--setup", str(choice), str(nx), str(ny), repr(eta), repr(t_end), str(tmp_path / "out.bin"),
            str(every)]
    ref = _run_context_sequence(tmp_path, nx, ny, dx, dy, eta, t_end, every, setup=name)
    init_b = _check_call_site(tmp_path, args, nx, ny, every, ref, True)
    host = setups.SETUPS[name](nx, ny, dx, dy)
    for q, k in enumerate(CONSERVED):
        a, b = init_b[q][2:-2, 2:-2], host[k][2:-2, 2:-2]
        assert np.abs(a - b).max() <= 1e-13 * np.abs(b).max(), (name, k)
