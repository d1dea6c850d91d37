"""GPU checks of the batched adjoint (mhd_batch_adjoint_* of the C ABI).

* Parity: every run's dJ/dU(initial) and dJ/d eta from one batched backward
  pass against a single-run context's adjoint (mhd_adjoint_*) of the same run
  -- the per-run arithmetic is the same; the CG dot products and the eta sums
  run in another order, so the bar is 1e-8 relative (max norm per field).
  Runs of uneven size mix the two adjoint solvers: Jacobi sweeps where the
  operator is close to the identity, CG for the run whose fixed step is far
  above the resistive limit.
* The recorded forward pass is the plain batched run (bitwise), and a second
  backward pass, which recomputes the last segment from its checkpoint, gives
  bitwise the same gradient as the first, which used the forward's stage data.
* Dot-product test of the batched gradient against central differences of the
  batched forward path: <lam0, v> vs <lamN, (F(U + e v) - F(U - e v)) / 2e>
  per run, with the 2R perturbed runs advanced as one batch (fixed CG
  iteration count, as in the single-run adjoint tests, so that the stopping
  test does not jump between the two sides).
"""
import numpy as np
import pytest

from helpers import CONSERVED, interior
from magnetohydrodynamics_simulator_amd import Batch, Context, setups

pytestmark = pytest.mark.gpu

FIXED_CG = dict(cg_tol=0.0, cg_max_iter=16)
TOL_PARITY = 1e-8
TOL_DOT = 1e-6

# (side, eta, dt or None for 0.8 x the CFL step at Courant number 0.4)
RUNS = [
    (32, 1e-2, None),
    (48, 3e-3, None),
    (32, 0.3, 2e-3),     # fixed step far above the resistive limit: CG adjoint solves
    (40, 1e-3, None),
]


def _say(msg):
    print(f"\n[case] {msg}", flush=True)


def _cases():
    out = []
    for k, (n, eta, dt) in enumerate(RUNS):
        amps, phases = setups.sausage_column_modes(8, 1e-2, seed=20 + k)
        st = setups.sausage_column(n, amps, phases)
        if dt is None:
            with Context(n, n, 1.0 / n, 1.0 / n, eta, cfl=0.4) as c:
                c.upload(*[st[f] for f in CONSERVED])
                dt = 0.8 * c.CFL_condition_checker()
        out.append((n, eta, dt, st))
    return out


def _smooth(n, seed, scale=1.0):
    rng = np.random.default_rng(seed)
    X, Y = np.meshgrid((np.arange(n + 4) - 2) / n, (np.arange(n + 4) - 2) / n)
    out = {}
    for f in CONSERVED:
        g = np.zeros_like(X)
        for _ in range(3):
            kx, ky = rng.integers(-2, 3, size=2)
            g += rng.standard_normal() * np.sin(2 * np.pi * (kx * X + ky * Y) + rng.uniform(0, 2 * np.pi))
        out[f] = scale * g
    return out


def _batch(cases, states=None, **opts):
    b = Batch([c[0] for c in cases], [c[0] for c in cases], [1.0 / c[0] for c in cases],
              [1.0 / c[0] for c in cases], [c[1] for c in cases], **opts)
    b.set_schedule(dt_fixed=[c[2] for c in cases])
    for r, c in enumerate(cases):
        b.upload(r, c[3] if states is None else states[r])
    return b


def _dot(a, b):
    return float(sum(np.sum(interior(a[k]) * interior(b[k])) for k in CONSERVED))


@pytest.mark.parametrize("every,opts", [(2, FIXED_CG), (10, dict(cg_tol=1e-12))],
                         ids=["every2-fixedcg", "every10-tol1e-12"])
def test_batch_adjoint_matches_single_run(every, opts):
    steps = 5
    cases = _cases()
    _say(f"batched adjoint of {len(cases)} sausage runs, {steps} steps, checkpoints every {every}, {opts}")
    lams = [_smooth(c[0], 100 + r) for r, c in enumerate(cases)]
    with _batch(cases, **opts) as b:
        b.adjoint_forward(steps, every)
        fin_b = [b.download(r) for r in range(len(cases))]
        for r in range(len(cases)):
            b.adjoint_set_lambda(r, lams[r])
        b.adjoint_backward()
        res_b = [b.adjoint_result(r) for r in range(len(cases))]
    with _batch(cases, **opts) as p:     # the plain batched run of the same schedule
        p.set_schedule(max_steps=steps)
        p.run()
        fin_p = [p.download(r) for r in range(len(cases))]
    for r, (n, eta, dt, st) in enumerate(cases):
        for f in CONSERVED:
            assert np.array_equal(interior(fin_b[r][f]), interior(fin_p[r][f])), (r, f)
        with Context(n, n, 1.0 / n, 1.0 / n, eta, **opts) as c:
            c.upload(*[st[f] for f in CONSERVED])
            c.adjoint_forward(dt, steps, every)
            fin_s = c.download()
            lam0_s, eb_s = c.adjoint_backward(lams[r])
        lam0_b, eb_b = res_b[r]
        for f in CONSERVED:
            ref = np.abs(interior(fin_s[f])).max()
            assert np.abs(interior(fin_b[r][f]) - interior(fin_s[f])).max() <= 1e-11 * ref, (r, f)
            scale = np.abs(interior(lam0_s[f])).max()
            err = np.abs(interior(lam0_b[f]) - interior(lam0_s[f])).max()
            assert err <= TOL_PARITY * scale, (r, f, err, scale)
        assert abs(eb_b - eb_s) <= TOL_PARITY * abs(eb_s), (r, eb_b, eb_s)


def test_batch_adjoint_repeat_recomputes_identically():
    steps, every = 5, 2
    cases = _cases()
    _say(f"batched adjoint twice, {steps} steps, checkpoints every {every}")
    lams = [_smooth(c[0], 200 + r) for r, c in enumerate(cases)]
    with _batch(cases, **FIXED_CG) as b:
        b.adjoint_forward(steps, every)
        out = []
        for _ in range(2):
            for r in range(len(cases)):
                b.adjoint_set_lambda(r, lams[r])
            b.adjoint_backward()
            out.append([b.adjoint_result(r) for r in range(len(cases))])
    for r in range(len(cases)):
        (l1, e1), (l2, e2) = out[0][r], out[1][r]
        assert e1 == e2, r
        for f in CONSERVED:
            assert np.array_equal(l1[f], l2[f]), (r, f)


def test_batch_adjoint_dot_product():
    steps, every, eps = 3, 2, 1e-6
    cases = _cases()
    _say(f"batched adjoint dot-product test, {steps} steps, eps {eps}")
    lams = [_smooth(c[0], 300 + r) for r, c in enumerate(cases)]
    vs = [_smooth(c[0], 400 + r, 1e-2) for r, c in enumerate(cases)]
    with _batch(cases, **FIXED_CG) as b:
        b.adjoint_forward(steps, every)
        for r in range(len(cases)):
            b.adjoint_set_lambda(r, lams[r])
        b.adjoint_backward()
        lam0 = [b.adjoint_result(r)[0] for r in range(len(cases))]
    pert_cases, pert_states = [], []
    for r, c in enumerate(cases):
        for sgn in (1, -1):
            pert_cases.append(c)
            pert_states.append({f: c[3][f] + sgn * eps * vs[r][f] for f in CONSERVED})
    with _batch(pert_cases, pert_states, **FIXED_CG) as p:
        p.set_schedule(max_steps=steps)
        p.run()
        outs = [p.download(r) for r in range(len(pert_cases))]
    for r in range(len(cases)):
        jv = {f: (outs[2 * r][f] - outs[2 * r + 1][f]) / (2 * eps) for f in CONSERVED}
        lhs, rhs = _dot(lam0[r], vs[r]), _dot(lams[r], jv)
        assert abs(lhs - rhs) <= TOL_DOT * abs(rhs), (r, lhs, rhs)


def test_batch_adjoint_argument_errors():
    cases = _cases()[:2]
    _say("batched adjoint argument errors")
    with _batch(cases) as b:
        with pytest.raises(RuntimeError):
            b.adjoint_backward()                       # nothing recorded
        b.set_schedule(dt_fixed=[cases[0][2], 0.0])    # run 1 on the CFL step
        with pytest.raises(RuntimeError):
            b.adjoint_forward(2, 1)
        b.set_schedule(dt_fixed=[c[2] for c in cases], max_steps=[1, 5])
        with pytest.raises(RuntimeError):              # run 0 stops after one step
            b.adjoint_forward(2, 1)
        with pytest.raises(RuntimeError):
            b.adjoint_result(0)
    eta0 = [(cases[0][0], 0.0, cases[0][2], cases[0][3])]
    with _batch(eta0) as b:
        with pytest.raises(RuntimeError):
            b.adjoint_forward(2, 1)


def test_batch_adjoint_non_square_runs_match_single_run():
    """Non-square runs (the fused adjoint march's strips and segments in
    both directions) against the single-run adjoint."""
    steps, every = 3, 2
    shapes = [(70, 24), (24, 70), (130, 16)]
    _say(f"batched adjoint of non-square runs {shapes}, {steps} steps")
    cases = []
    for r, (nx, ny) in enumerate(shapes):
        n = max(nx, ny)
        amps, phases = setups.sausage_column_modes(8, 1e-2, seed=40 + r)
        full = setups.sausage_column(n, amps, phases)
        st = {k: np.ascontiguousarray(full[k][: ny + 4, : nx + 4]) for k in CONSERVED}
        cases.append((nx, ny, 1.0 / n, 1e-2, st))
    lams = [_smooth(max(nx, ny), 500 + r) for r, (nx, ny, *_rest) in enumerate(cases)]
    lams = [{k: np.ascontiguousarray(l[k][: ny + 4, : nx + 4]) for k in CONSERVED}
            for l, (nx, ny, *_rest) in zip(lams, cases)]
    dt = 2e-4
    b = Batch([c[0] for c in cases], [c[1] for c in cases], [c[2] for c in cases], [c[2] for c in cases],
              [c[3] for c in cases], **FIXED_CG)
    with b:
        b.set_schedule(dt_fixed=[dt] * len(cases))
        for r, c in enumerate(cases):
            b.upload(r, c[4])
        b.adjoint_forward(steps, every)
        for r in range(len(cases)):
            b.adjoint_set_lambda(r, lams[r])
        b.adjoint_backward()
        res = [b.adjoint_result(r) for r in range(len(cases))]
    for r, (nx, ny, h, eta, st) in enumerate(cases):
        with Context(nx, ny, h, h, eta, **FIXED_CG) as c:
            c.upload(*[st[k] for k in CONSERVED])
            c.adjoint_forward(dt, steps, every)
            lam0, eb = c.adjoint_backward(lams[r])
        for k in CONSERVED:
            scale = np.abs(interior(lam0[k])).max()
            err = np.abs(interior(res[r][0][k]) - interior(lam0[k])).max()
            assert err <= TOL_PARITY * scale, (r, k, err, scale)
        assert abs(res[r][1] - eb) <= TOL_PARITY * abs(eb), (r, res[r][1], eb)


def test_config5_full_size():
    """The batched-differentiation workload at its full size: 64 seeded
    sausage-column runs (sides 256..1024, 33.5 M cells, eta 1e-4..1e-2), each at
    0.8 x its CFL step, 40 steps, checkpoints every 10, loss = each run's
    window magnetic energy.
      * the recorded forward is the plain batched forward bit for bit;
      * the two smallest runs' dJ/dU(initial) and dJ/d eta against single-run
        contexts: 1e-8 relative;
      * three runs' gradients against central differences of the batched
        forward (a batch of their 6 perturbed states, 16 fixed CG
        iterations): <lam0, v> vs <lamN, (F(U + e v) - F(U - e v)) / 2e>,
        1e-6 relative."""
    steps, every, eps = 40, 10, 1e-5     # state step 1e-8 (noise grows as 1/eps below it: 3e-6 at 1e-10)
    members = setups.sausage_ensemble(64, seed=0)
    ns = [m[0] for m in members]
    etas = [m[1] for m in members]
    modes = [setups.sausage_column_modes(8, 1e-2, seed=s) for _, _, s in members]
    _say(f"config 5: {len(members)} sausage runs, sides {min(ns)}..{max(ns)}, {sum(n * n for n in ns)} cells, "
         f"{steps} steps, checkpoints every {every}")

    def make():
        b = Batch(ns, ns, [1.0 / n for n in ns], [1.0 / n for n in ns], etas, cfl=0.4, cg_tol=1e-10)
        for r, md in enumerate(modes):
            b.init_column(r, md[0], md[1], True)
        return b

    order = sorted(range(len(members)), key=lambda r: (ns[r], r))
    small = order[:2]
    fd_runs = order[:2] + [next(r for r in order if ns[r] == 512)]
    with make() as b:
        b.set_schedule(max_steps=1)
        b.run()
        dts = [0.8 * b.run_info(r)[2] for r in range(len(members))]
    with make() as p:
        p.set_schedule(max_steps=steps, dt_fixed=dts)
        p.run()
        plain = {r: p.download(r) for r in set(small + fd_runs)}
    lams, res = {}, {}
    with make() as b:
        b.set_schedule(max_steps=steps, dt_fixed=dts)
        b.adjoint_forward(steps, every)
        for r in range(len(members)):
            fin = b.download(r)
            if r in plain:
                for f in CONSERVED:
                    assert np.array_equal(interior(fin[f]), interior(plain[r][f])), (r, f)
            _, lam = setups.window_magnetic_energy(fin, ns[r], 1.0 / ns[r])
            lams[r] = lam
            b.adjoint_set_lambda(r, lam)
        b.adjoint_backward()
        for r in set(small + fd_runs):
            res[r] = b.adjoint_result(r)
    print(f"[case] recorded forward == plain batched forward on runs {sorted(plain)}", flush=True)
    for r in small:
        n = ns[r]
        with Context(n, n, 1.0 / n, 1.0 / n, etas[r], cfl=0.4, cg_tol=1e-10) as c:
            c.init_column(modes[r][0], modes[r][1], True)
            c.adjoint_forward(dts[r], steps, every)
            lam0_s, eb_s = c.adjoint_backward(lams[r])
        lam0_b, eb_b = res[r]
        worst = max(np.abs(interior(lam0_b[f]) - interior(lam0_s[f])).max() / np.abs(interior(lam0_s[f])).max()
                    for f in CONSERVED)
        print(f"[case] run {r} ({n}^2, eta {etas[r]:.3e}): lam0 max rel diff {worst:.2e}, "
              f"eta_bar {eb_b:.12e} vs {eb_s:.12e}", flush=True)
        assert worst <= TOL_PARITY and abs(eb_b - eb_s) <= TOL_PARITY * abs(eb_s), (r, worst, eb_b, eb_s)
    # central differences of the batched forward for three runs
    vs = {r: _smooth(ns[r], 500 + r, 1e-3) for r in fd_runs}
    pert_ns, pert_etas, pert_dts, states = [], [], [], []
    for r in fd_runs:
        st = setups.sausage_column(ns[r], modes[r][0], modes[r][1])
        for sgn in (1, -1):
            pert_ns.append(ns[r])
            pert_etas.append(etas[r])
            pert_dts.append(dts[r])
            states.append({f: st[f] + sgn * eps * vs[r][f] for f in CONSERVED})
    with Batch(pert_ns, pert_ns, [1.0 / n for n in pert_ns], [1.0 / n for n in pert_ns], pert_etas,
               **FIXED_CG) as q:
        q.set_schedule(max_steps=steps, dt_fixed=pert_dts)
        for k, st in enumerate(states):
            q.upload(k, st)
        q.run()
        outs = [q.download(k) for k in range(len(states))]
    for i, r in enumerate(fd_runs):
        jv = {f: (outs[2 * i][f] - outs[2 * i + 1][f]) / (2 * eps) for f in CONSERVED}
        lhs, rhs = _dot(res[r][0], vs[r]), _dot(lams[r], jv)
        print(f"[case] run {r} ({ns[r]}^2): <lam0, v> {lhs:.12e} vs differences {rhs:.12e}, "
              f"rel {abs(lhs - rhs) / abs(rhs):.2e}", flush=True)
        assert abs(lhs - rhs) <= TOL_DOT * abs(rhs), (r, lhs, rhs)
