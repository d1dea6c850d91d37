"""GPU checks of the device-resident driver loop (mhd_run_resident).

The resident loop takes the driver's per-step decisions on the device (CFL
dt, the t_end clamp, the operator rebuild rule, CG termination) and must give
exactly what mhd_run gives -- the same kernels on the same operands, so the
bar is bitwise: every conserved field, every dt, every CG iteration count and
the final time.  mhd_run itself is held to the oracle by test_gpu_parity.py.

Cases: the adaptive CG budget; fixed budgets of 0 and 1 launched iterations,
which halt the loop in (nearly) every solve and send those steps through the
host's resume path, with group sizes 1, 3 and 16; solves of 18-22 iterations
halted and resumed mid-solve; pairs of steps replayed as
captured graphs (every replay of a group checked through the final state); the t_end clamp; a constant
dt; eta = 0 (no implicit solve); a non-square grid; the snapshot cadence
(frame numbers, the state at each frame, files in the reference's format);
and the error paths (non-finite wave speed, the CG iteration limit with
cg_fail_is_error, an exception in the callback).
"""
import os

import numpy as np
import pytest

from helpers import CONSERVED, ic_arrays, make_gpu

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"\n[case] {msg}", flush=True)


def _plain(ic, nx, ny, dx, dy, eta, t_end, max_steps, dt=0.0, **kw):
    with make_gpu(ic, nx, ny, dx, dy, eta, **kw) as ctx:
        n, dts, its, t = ctx.run(t_end, dt_override=dt, max_steps=max_steps)
        return n, dts, its, t, ctx.download()


def _resident(ic, nx, ny, dx, dy, eta, t_end, max_steps, dt=0.0, tune=None, **kw):
    with make_gpu(ic, nx, ny, dx, dy, eta, **kw) as ctx:
        if tune:
            ctx.resident_tune(*tune)
        n, dts, its, t = ctx.run_resident(t_end, dt_override=dt, max_steps=max_steps)
        return n, dts, its, t, ctx.download()


def _same(a, b):
    na, dta, ita, ta, fa = a
    nb, dtb, itb, tb, fb = b
    assert na == nb, (na, nb)
    assert np.array_equal(dta, dtb)
    assert np.array_equal(ita, itb), (ita, itb)
    assert ta == tb, (ta, tb)
    for k in CONSERVED:
        assert np.array_equal(fa[k], fb[k]), k


@pytest.mark.parametrize("tune", [None, (16, 1, 0), (3, 1, 1), (1, 0, -1), (5, 0, 0), (16, 0, -1, True),
                                  (5, 0, 0, True), (4, 1, 1, True), (16, 0, -1, 2), (5, 0, 0, 2)],
                         ids=["adaptive", "fixed0-g16", "fixed1-g3", "margin0-g1", "fixed0-g5", "graph-g16",
                              "graph-fixed0-g5", "graph-fixed1-g4", "groupgraph-g16", "groupgraph-fixed0-g5"])
def test_resident_matches_run(tune):
    n, eta = 96, 1e-2
    _say(f"resident vs mhd_run, OT {n}^2 eta={eta}, 25 steps, tune={tune}")
    ic, dx, dy = ic_arrays("orszag_tang", n, n)
    ref = _plain(ic, n, n, dx, dy, eta, 10.0, 25, tol=1e-10, cfl=0.4)
    got = _resident(ic, n, n, dx, dy, eta, 10.0, 25, tune=tune, tol=1e-10, cfl=0.4)
    assert ref[2].min() > 0          # the solves iterate
    _same(got, ref)


@pytest.mark.parametrize("tune", [None, (3, 0, 0), (4, 1, 1, True)], ids=["adaptive", "fixed0-g3", "graph-fixed1-g4"])
def test_long_solves(tune):
    """Solves of many iterations: resistivity-limited steps at tol 1e-25 (the
    recursive residual keeps falling) take 18-22 iterations per solve, so the
    fixed budgets halt the loop deep inside a solve and the host resumes it
    there.  Resident loop = mhd_run bit for bit, and mhd_run = the oracle's
    loop (same time steps, iteration counts within one, relative L1 <= 1e-10)."""
    from helpers import interior, make_oracle, rel_l1
    n, eta, tol, steps = 64, 0.5, 1e-25, 4
    _say(f"long solves: resident vs mhd_run vs oracle, OT {n}^2 eta={eta} tol={tol}, tune={tune}")
    ic, dx, dy = ic_arrays("orszag_tang", n, n)
    ref = _plain(ic, n, n, dx, dy, eta, 10.0, steps, tol=tol, cfl=0.4)
    assert ref[2].min() > 16, ref[2]
    got = _resident(ic, n, n, dx, dy, eta, 10.0, steps, tune=tune, tol=tol, cfl=0.4)
    _same(got, ref)
    if tune is None:
        sim = make_oracle(ic, n, n, dx, dy, eta, tol=tol)
        nr, dts_r, its_r, _ = sim.run(1e9, cfl=0.4, max_steps=steps, snapshot_every=0)
        assert nr == steps and np.allclose(ref[1], dts_r, rtol=1e-12, atol=0)
        assert np.abs(ref[2] - its_r).max() <= 1, (ref[2], its_r)
        for k in CONSERVED:
            e = rel_l1(interior(ref[4][k]), sim.interior(k))
            assert e <= 1e-10, (k, e)
        sim.close()


def test_resident_group_graph_chunks():
    """A group longer than one captured graph (graph = 2 captures at most 64
    steps): a 140-step group runs as graphs of 64, 64 and 12 steps, bitwise
    mhd_run."""
    n, eta, steps = 48, 1e-2, 140
    _say(f"resident group of {steps} steps as chunked graphs, OT {n}^2 eta={eta}")
    ic, dx, dy = ic_arrays("orszag_tang", n, n)
    ref = _plain(ic, n, n, dx, dy, eta, 10.0, steps, tol=1e-10, cfl=0.4)
    got = _resident(ic, n, n, dx, dy, eta, 10.0, steps, tune=(200, 1, -1, 2), tol=1e-10, cfl=0.4)
    _same(got, ref)


def test_resident_t_end_clamp_and_eta0():
    n = 64
    ic, dx, dy = ic_arrays("orszag_tang", n, n)
    for eta in (0.0, 3e-3):
        _say(f"resident to t_end=0.05 (clamped last step), OT {n}^2 eta={eta}")
        ref = _plain(ic, n, n, dx, dy, eta, 0.05, -1, cfl=0.4, tol=1e-12)
        got = _resident(ic, n, n, dx, dy, eta, 0.05, -1, cfl=0.4, tol=1e-12)
        assert ref[3] == 0.05 and ref[0] > 2
        assert ref[1][-1] < ref[1][-2]   # the clamp shortened the last step
        _same(got, ref)


def test_resident_fixed_dt_non_square():
    nx, ny, eta = 90, 40, 5e-3
    _say(f"resident with dt_override, {nx}x{ny} eta={eta}")
    ic, dx, dy = ic_arrays("orszag_tang", nx, ny, seed=3, amp=1e-3)
    ref = _plain(ic, nx, ny, dx, dy, eta, 1.0, 12, dt=2e-3, cfl=0.4, tol=1e-11)
    got = _resident(ic, nx, ny, dx, dy, eta, 1.0, 12, dt=2e-3, tune=(4, 1, -1), cfl=0.4, tol=1e-11)
    _same(got, ref)
    got = _resident(ic, nx, ny, dx, dy, eta, 1.0, 12, dt=2e-3, tune=(4, 0, -1, True), cfl=0.4, tol=1e-11)
    assert np.all(ref[1] == 2e-3)
    _same(got, ref)


@pytest.mark.parametrize("shape", [(300, 48), (256, 70), (40, 300)], ids=["300x48", "256x70", "40x300"])
def test_resident_tile_edges(shape):
    """The final recombination forms the next step's CFL partials itself: grids
    with a partial last column of tiles, an exact multiple of the tile width,
    and tall grids (many tile rows, the wrap of the top tile)."""
    nx, ny = shape
    _say(f"resident tile edges {nx}x{ny}, eta=2e-3, 8 CFL steps")
    ic, dx, dy = ic_arrays("orszag_tang", nx, ny, seed=5, amp=1e-3)
    ref = _plain(ic, nx, ny, dx, dy, 2e-3, 10.0, 8, cfl=0.4, tol=1e-10)
    got = _resident(ic, nx, ny, dx, dy, 2e-3, 10.0, 8, cfl=0.4, tol=1e-10)
    _same(got, ref)


def test_resident_snapshot_cadence(tmp_path):
    n, eta, every, steps = 64, 1e-3, 5, 12
    _say(f"resident snapshots every {every}, {steps} steps")
    ic, dx, dy = ic_arrays("orszag_tang", n, n)
    frames, states = [], {}

    def on_snap(ctx, frame, t, dt):
        frames.append((frame, t, dt))
        states[frame] = ctx.download()
        ctx.write_snapshot(os.fsdecode(tmp_path / f"snapshot_{frame:06d}.dat"), every, dt)

    with make_gpu(ic, n, n, dx, dy, eta, cfl=0.4, tol=1e-10) as ctx:
        dt0 = ctx.CFL_condition_checker()
        ns, dts, _, t = ctx.run_resident(10.0, max_steps=steps, snapshot_every=every, on_snapshot=on_snap)
    assert ns == steps
    assert [f for f, _, _ in frames] == [0, 5, 10]
    assert frames[0][1] == 0.0 and frames[0][2] == dt0
    for f, tt, dt in frames[1:]:
        assert dt == dts[f - 1] and tt == pytest.approx(float(np.sum(dts[:f])), rel=1e-15)
    for f in (0, 5, 10):
        ref = _plain(ic, n, n, dx, dy, eta, 10.0, f, cfl=0.4, tol=1e-10)
        for k in CONSERVED:
            assert np.array_equal(states[f][k], ref[4][k]), (f, k)
        head = np.fromfile(tmp_path / f"snapshot_{f:06d}.dat", dtype=np.float64, count=9)
        assert list(head[:4]) == [n, n, 2, every]


@pytest.mark.parametrize("case", ["periodic", "outflow", "graph", "groupgraph"])
def test_async_snapshot_files(tmp_path, case):
    """mhd_run_resident_files writes, while the loop runs on, byte for byte
    the files a mhd_run_resident callback writes with mhd_write_snapshot at
    the same frames (before the first step and every 3rd step), and leaves
    the same run (state, time steps, CG counts)."""
    nx, ny, eta, every, steps = 64, 48, 1e-2, 3, 14
    bc = ("outflow", "periodic") if case == "outflow" else ("periodic", "periodic")
    _say(f"asynchronous snapshot files, {nx}x{ny} {case}, every {every}, {steps} steps")
    ic, dx, dy = ic_arrays("orszag_tang", nx, ny, seed=3, amp=1e-2)
    sync_dir, async_dir = tmp_path / "sync", tmp_path / "async"
    sync_dir.mkdir()
    async_dir.mkdir()

    def on_snap(ctx, frame, t, dt):
        ctx.write_snapshot(os.fsdecode(sync_dir / f"snapshot_{frame:06d}.dat"), every, dt)

    with make_gpu(ic, nx, ny, dx, dy, eta, cfl=0.4, tol=1e-10, bc=bc) as ctx:
        if case in ("graph", "groupgraph"):
            ctx.resident_tune(4, 0, -1, 2 if case == "groupgraph" else 1)
        ref = ctx.run_resident(10.0, max_steps=steps, snapshot_every=every, on_snapshot=on_snap)
        ref_state = ctx.download()
    with make_gpu(ic, nx, ny, dx, dy, eta, cfl=0.4, tol=1e-10, bc=bc) as ctx:
        if case in ("graph", "groupgraph"):
            ctx.resident_tune(4, 0, -1, 2 if case == "groupgraph" else 1)
        got = ctx.run_resident_files(10.0, every, os.fsdecode(async_dir / "snapshot_%06d.dat"), max_steps=steps)
        got_state = ctx.download()
    assert got[0] == ref[0] == steps
    assert np.array_equal(got[1], ref[1]) and np.array_equal(got[2], ref[2]) and got[3] == ref[3]
    for k in CONSERVED:
        assert np.array_equal(got_state[k], ref_state[k]), k
    names = sorted(p.name for p in sync_dir.iterdir())
    assert names == [f"snapshot_{f:06d}.dat" for f in range(0, steps + 1, every)]
    assert sorted(p.name for p in async_dir.iterdir()) == names
    for name in names:
        assert (async_dir / name).read_bytes() == (sync_dir / name).read_bytes(), name


def test_async_snapshot_errors(tmp_path):
    n = 32
    ic, dx, dy = ic_arrays("orszag_tang", n, n)
    _say("asynchronous snapshot files: error paths")
    with make_gpu(ic, n, n, dx, dy, 1e-3, cfl=0.4) as ctx:
        for fmt in ("snap_%s.dat", "snap_%d_%d.dat", "snap.dat"):
            with pytest.raises(RuntimeError, match="path format"):
                ctx.run_resident_files(1.0, 2, os.fsdecode(tmp_path / fmt), max_steps=2)
        with pytest.raises(RuntimeError, match="cannot open"):
            ctx.run_resident_files(1.0, 2, os.fsdecode(tmp_path / "missing_dir" / "snap_%04d.dat"), max_steps=6)
        # the context still runs
        ns, _, _, _ = ctx.run_resident(1.0, max_steps=2)
        assert ns == 2


def test_resident_errors():
    n = 32
    ic, dx, dy = ic_arrays("orszag_tang", n, n)
    _say("resident error paths")
    bad = {k: v.copy() for k, v in ic.items()}
    bad["E"][10, 10] = np.nan
    with make_gpu(bad, n, n, dx, dy, 1e-3, cfl=0.4) as ctx:
        with pytest.raises(RuntimeError, match="non-finite"):
            ctx.run_resident(1.0, max_steps=3)
    with make_gpu(ic, n, n, dx, dy, 1e-1, cfl=0.4, cg_max_iter=1, cg_fail_is_error=True) as ctx:
        with pytest.raises(RuntimeError, match="iteration limit"):
            ctx.run_resident(1.0, max_steps=3)

    def boom(ctx, frame, t, dt):
        if frame == 2:
            raise KeyError("stop here")

    with make_gpu(ic, n, n, dx, dy, 1e-3, cfl=0.4) as ctx:
        with pytest.raises(KeyError):
            ctx.run_resident(1.0, max_steps=6, snapshot_every=2, on_snapshot=boom)
    with make_gpu(ic, n, n, dx, dy, 1e-3, cfl=0.4) as ctx:
        with pytest.raises(RuntimeError):
            ctx.resident_tune(0, 1, -1)
