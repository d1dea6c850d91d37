"""GPU checks of the snapshot and restart files (mhd_write_snapshot,
mhd_write_restart, mhd_read_restart).  The snapshot is read the way the
reference's reader does it (scripts/display_functions.py:3-22: 9 header
doubles, then seven padded arrays rho, E, p, vx, vy, bx, by) and compared
with mhd_download of the same state; a run continued from a restart file
equals the uninterrupted run."""
import numpy as np
import pytest

from helpers import CONSERVED, ic_arrays
from magnetohydrodynamics_simulator_amd import Context

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"\n[case] {msg}", flush=True)


def test_snapshot_matches_download_and_reader_layout(tmp_path):
    nx, ny, eta = 40, 24, 1e-2
    _say(f"snapshot file, OT {nx}x{ny}")
    ic, dx, dy = ic_arrays("orszag_tang", nx, ny, seed=1, amp=1e-3)
    path = tmp_path / "snapshot_000003.dat"
    with Context(nx, ny, dx, dy, eta) as c:
        c.upload(*[ic[k] for k in CONSERVED])
        c.run(1e9, max_steps=3)
        c.write_snapshot(str(path), 20, 1.25e-3)
        full = c.download(primitives=True)
        rms, mx = c.calculate_current_div_B()
    data = np.fromfile(path, dtype=np.float64)
    nxt, nyt, ng = nx + 4, ny + 4, 2
    assert data.size == 9 + 7 * nxt * nyt
    hdr = data[:9]
    assert list(hdr[:4]) == [nx, ny, ng, 20]
    assert hdr[4] == dx and hdr[5] == dy and hdr[6] == 1.25e-3
    assert hdr[7] == rms and hdr[8] == mx
    fields = data[9:].reshape(-1, nyt, nxt)
    for q, k in enumerate(("rho", "E", "p", "vx", "vy", "bxc", "byc")):
        assert np.array_equal(fields[q], full[k]), k


def test_restart_continues_the_run_exactly(tmp_path):
    n, eta = 48, 1e-3
    _say(f"restart file, OT {n}x{n}, 4 + 4 steps vs 8 steps with a fixed step")
    ic, dx, dy = ic_arrays("orszag_tang", n, n, seed=2, amp=1e-3)
    path = str(tmp_path / "state.rst")
    with Context(n, n, dx, dx, eta) as c:
        c.upload(*[ic[k] for k in CONSERVED])
        dt = 0.5 * c.CFL_condition_checker()
        c.run(1e9, dt_override=dt, max_steps=8)
        ref = c.download()
    with Context(n, n, dx, dx, eta) as c:
        c.upload(*[ic[k] for k in CONSERVED])
        c.run(1e9, dt_override=dt, max_steps=4)
        c.write_restart(path, 4 * dt)
    with Context(n, n, dx, dx, eta) as c:
        t = c.read_restart(path)
        assert t == 4 * dt
        c.run(1e9, dt_override=dt, max_steps=4)
        out = c.download()
    for k in CONSERVED:
        assert np.array_equal(out[k], ref[k]), k
    with Context(n + 2, n, dx, dx, eta) as c:
        with pytest.raises(RuntimeError, match="another grid"):
            c.read_restart(path)
