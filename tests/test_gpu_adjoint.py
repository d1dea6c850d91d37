"""GPU checks of the discrete adjoint (mhd_adjoint_* and
mhd_explicit_rhs_vjp of the C ABI) by dot-product tests against central
finite differences of the GPU forward path on the same inputs:
  <L'(U)^T lam, v>       vs  <lam, (L(U + e v) - L(U - e v)) / 2e>
  <lam0, v>              vs  <lamN, (F(U0 + e v) - F(U0 - e v)) / 2e>
  eta_bar                vs  (J(eta + h) - J(eta - h)) / 2h
with F = N IMEX steps of constant dt and J = <lamN, U_N>.  States are smooth
(no limiter kink lies within e of any cell); the CG tolerance is the
reference default (DBL_EPSILON) so that the solves do not add noise to the
differences.  Tolerance: 1e-6 relative (the differentiation bar).  Multi-step runs use
smooth (low-mode) directions: with grid-scale random directions the
central differences straddle minmod kinks somewhere in the run and the
comparison measures that noise (about 1e-4 at 4 steps) rather than the
adjoint, which agrees to 1e-9 per step."""
import numpy as np
import pytest

from helpers import CONSERVED, interior
from magnetohydrodynamics_simulator_amd import Context, setups

pytestmark = pytest.mark.gpu

TOL = 1e-6


def _say(msg):
    print(f"\n[case] {msg}", flush=True)


def _state(n, seed):
    """Fourier-perturbed member with smooth non-uniform density and pressure:
    a uniform field puts every minmod of that variable at its kink
    (both slopes zero), where central differences do not measure a
    derivative."""
    ic = dict(setups.ensemble_state(n, seed))
    dx = setups.grid_spacing(n)
    X, Y = np.meshgrid((np.arange(n + 4) - 1.5) * dx, (np.arange(n + 4) - 1.5) * dx)
    drho = 0.15 * np.sin(X + 2 * Y + 0.4) + 0.1 * np.cos(3 * X - Y + 1.1)
    dpr = 0.2 * np.cos(2 * X + Y + 0.3) + 0.05 * np.sin(X - 3 * Y + 2.0)
    ic["rho"] = ic["rho"] * (1 + drho)
    ic["mx"] = ic["mx"] * (1 + drho)
    ic["my"] = ic["my"] * (1 + drho)
    ic["E"] = ic["E"] + dpr / (5.0 / 3.0 - 1.0)
    return ic


def _rand_fields(n, seed, scale=1.0):
    rng = np.random.default_rng(seed)
    return [scale * rng.standard_normal((n + 4, n + 4)) for _ in range(6)]


def _smooth_fields(n, seed, scale=1.0):
    """Low-mode (|k| <= 3) periodic fields: a perturbation of this shape moves
    the minmod slope differences by O(dx^2) of its size, so the central
    differences of a multi-step run rarely straddle a limiter kink."""
    rng = np.random.default_rng(seed)
    dx = setups.grid_spacing(n)
    X, Y = np.meshgrid((np.arange(n + 4) - 2) * dx, (np.arange(n + 4) - 2) * dx)
    out = []
    for _ in range(6):
        f = np.zeros_like(X)
        for _ in range(4):
            kx, ky = rng.integers(-3, 4, size=2)
            f += rng.standard_normal() * np.sin(kx * X + ky * Y + rng.uniform(0, 2 * np.pi))
        out.append(scale * f)
    return out


def _dot(a, b):
    return float(sum(np.sum(interior(x) * interior(y)) for x, y in zip(a, b)))


def _ctx(ic, n, eta):
    dx = setups.grid_spacing(n)
    c = Context(n, n, dx, dx, eta)            # default CG tolerance: DBL_EPSILON
    c.upload(*[ic[k] for k in CONSERVED])
    return c


def test_explicit_rhs_vjp_dot_product():
    _say("explicit_rhs_vjp dot product, 40x40 ensemble member")
    n = 40
    ic = _state(n, 7)
    lam = _rand_fields(n, 1)
    v = _rand_fields(n, 2, 1e-3)
    with _ctx(ic, n, 0.0) as c:
        adj = c.explicit_rhs_vjp(lam)
    eps = 1e-4
    outs = []
    for sgn in (1, -1):
        st = {k: ic[k] + sgn * eps * v[q] for q, k in enumerate(CONSERVED)}
        with _ctx(st, n, 0.0) as c:
            outs.append(c.explicit_rhs())
    jv = [(a - b) / (2 * eps) for a, b in zip(*outs)]
    lhs, rhs = _dot(adj, v), _dot(lam, jv)
    assert abs(lhs - rhs) <= TOL * abs(rhs), (lhs, rhs)


@pytest.mark.parametrize("nx,ny,seed", [(40, 40, 7), (70, 44, 8), (130, 9, 9), (5, 6, 10)])
def test_fused_rhs_vjp_matches_unfused(nx, ny, seed):
    """The fused adjoint march against the four unfused gather kernels it
    replaces (the same per-face formulas and summation order): equal to
    round-off, on grids narrower than one strip, wider than two, and with
    segments of a few rows."""
    _say(f"fused vs unfused L'^T lam, {nx}x{ny}")
    n = max(nx, ny)
    ic = _state(n, seed)
    st = {k: np.ascontiguousarray(ic[k][: ny + 4, : nx + 4]) for k in CONSERVED}
    rng = np.random.default_rng(seed)
    lam = [rng.standard_normal((ny + 4, nx + 4)) for _ in range(6)]
    dx = setups.grid_spacing(n)
    with Context(nx, ny, dx, dx, 0.0) as c:
        c.upload(*[st[k] for k in CONSERVED])
        fused = c.explicit_rhs_vjp(lam)
        ref = c.explicit_rhs_vjp(lam, fused=False)
    for q, k in enumerate(CONSERVED):
        a, b = interior(fused[q]), interior(ref[q])
        assert np.abs(a - b).max() <= 1e-13 * np.abs(b).max(), (k, np.abs(a - b).max(), np.abs(b).max())


def _run(ic, n, eta, dt, steps):
    with _ctx(ic, n, eta) as c:
        c.run(1e9, dt_override=dt, max_steps=steps)
        return c.download()


@pytest.mark.parametrize("steps,every", [(1, 1), (5, 2), (4, 10)])
def test_adjoint_run_dot_product(steps, every):
    _say(f"adjoint of {steps} steps, checkpoints every {every}, 32x32, eta 2e-2")
    n, eta, dt = 32, 2e-2, 4e-3
    ic = _state(n, 11)
    lamN = dict(zip(CONSERVED, _smooth_fields(n, 3)))
    v = _smooth_fields(n, 4, 1e-3)
    with _ctx(ic, n, eta) as c:
        c.adjoint_forward(dt, steps, every)
        fin = c.download()
        lam0, eta_bar = c.adjoint_backward(lamN)
    ref = _run(ic, n, eta, dt, steps)
    for k in CONSERVED:       # the recorded forward is the plain forward
        assert np.array_equal(interior(fin[k]), interior(ref[k])), k
    eps = 1e-4
    outs = []
    for sgn in (1, -1):
        st = {k: ic[k] + sgn * eps * v[q] for q, k in enumerate(CONSERVED)}
        outs.append(_run(st, n, eta, dt, steps))
    jv = [(outs[0][k] - outs[1][k]) / (2 * eps) for k in CONSERVED]
    lhs = _dot([lam0[k] for k in CONSERVED], v)
    rhs = _dot([lamN[k] for k in CONSERVED], jv)
    assert abs(lhs - rhs) <= TOL * abs(rhs), (lhs, rhs)
    # eta derivative of J = <lamN, U_N>
    h = 1e-4 * eta
    jp = _dot([lamN[k] for k in CONSERVED], [_run(ic, n, eta + h, dt, steps)[k] for k in CONSERVED])
    jm = _dot([lamN[k] for k in CONSERVED], [_run(ic, n, eta - h, dt, steps)[k] for k in CONSERVED])
    fd = (jp - jm) / (2 * h)
    assert abs(eta_bar - fd) <= 1e-5 * abs(fd), (eta_bar, fd)


# Central differences need a forward map without jumps; the CG stopping test
# is one (near the tolerance a perturbation adds or drops an iteration), so
# the gradient checks run forward and adjoint with a fixed iteration count.
FIXED_CG = dict(cg_tol=0.0, cg_max_iter=16)


def _kink_loss(n, amps, phases, eta, dt, steps):
    st = setups.kink_column(n, amps, phases)
    with Context(n, n, 1.0 / n, 1.0 / n, eta, **FIXED_CG) as c:
        c.upload(*[st[k] for k in CONSERVED])
        c.run(1e9, dt_override=dt, max_steps=steps)
        fin = c.download()
    return setups.window_magnetic_energy(fin, n, 1.0 / n)[0]


def test_kink_column_gradient_unfrozen_differences():
    """Reduced differentiation workload: kink column 64x64, 8 displacement
    modes, eta 1e-3, 12 steps of frozen dt, checkpoints every 5; the loss is
    the magnetic energy of the central window.  Gradient with respect to the
    amplitudes (lam0 contracted with the analytic d(state)/d(a_k)) and eta
    against central differences of the plain GPU forward -- the map with its
    branch choices free to switch.  Its differences cross minmod / upwind
    kinks, so the amplitude comparison resolves only about 1e-5 here (bar
    1e-4 on the norm; the 1e-6 differentiation bar is checked per mode on
    the branch-frozen map in test_gpu_frozen_fd.py); eta, whose differences
    cross no kink, is held to 1e-6."""
    n, eta, steps = 64, 1e-3, 12
    _say(f"kink column {n}^2 gradient, 8 modes, {steps} steps")
    amps, phases = setups.kink_column_modes(8, 1e-3, seed=0)
    st = setups.kink_column(n, amps, phases)
    with Context(n, n, 1.0 / n, 1.0 / n, eta, cfl=0.4, **FIXED_CG) as c:
        c.upload(*[st[k] for k in CONSERVED])
        dt = 0.8 * c.CFL_condition_checker()
        c.adjoint_forward(dt, steps, 5)
        fin = c.download()
        loss, lam = setups.window_magnetic_energy(fin, n, 1.0 / n)
        lam0, eta_bar = c.adjoint_backward(lam)
    grad = setups.kink_column_gradient(lam0, setups.kink_column_dstate(n, amps, phases))
    assert abs(loss - _kink_loss(n, amps, phases, eta, dt, steps)) <= 1e-13 * loss
    # the upwind E_z selection of the CT step (solvers.hpp:326-333) switches
    # discontinuously where a corner velocity changes sign, and over a dozen
    # steps a difference step can cross such a jump (more often the larger
    # it is); below 1e-8 the differences' own rounding grows (about 1e-9 / h).
    # Per mode: the differences at h = 1e-7 and 3e-8; where they disagree by
    # more than 1e-5 a jump was crossed and the smaller step is kept.
    fd = np.zeros_like(grad)
    for k in range(len(amps)):
        vals = []
        for h in (1e-7, 3e-8):
            e = np.zeros_like(amps)
            e[k] = h
            vals.append((_kink_loss(n, amps + e, phases, eta, dt, steps) -
                         _kink_loss(n, amps - e, phases, eta, dt, steps)) / (2 * h))
        fd[k] = 0.5 * (vals[0] + vals[1]) if abs(vals[0] - vals[1]) <= 1e-5 * abs(vals[1]) else vals[1]
    err = np.linalg.norm(grad - fd) / np.linalg.norm(fd)
    print(f"gradient norm {np.linalg.norm(fd):.6e}, relative error {err:.3e}", flush=True)
    print("per-mode relative error", np.abs(grad - fd) / np.abs(fd), flush=True)
    # the differences themselves limit this comparison (truncation and
    # limiter kinks at larger steps, rounding at smaller ones); the adjoint's
    # own accuracy is pinned by the dot-product tests above and by eta below
    assert err <= 1e-4, (grad, fd)
    he = 1e-4 * eta
    fde = (_kink_loss(n, amps, phases, eta + he, dt, steps) - _kink_loss(n, amps, phases, eta - he, dt, steps)) / (2 * he)
    print(f"eta: adjoint {eta_bar:.10e} differences {fde:.10e}", flush=True)
    assert abs(eta_bar - fde) <= TOL * abs(fde), (eta_bar, fde)


def test_adjoint_repeat_replays_the_same_sweep():
    """A second forward/backward of the same configuration on one context
    (the backward sweep is then replayed from its captured graph) gives the
    same adjoint bit for bit, and leaves the final state in place."""
    _say("adjoint repeated on one context, 32x32, 6 steps")
    n, eta, dt, steps = 32, 2e-2, 4e-3, 6
    ic = _state(n, 5)
    lamN = dict(zip(CONSERVED, _smooth_fields(n, 6)))
    res = []
    with _ctx(ic, n, eta) as c:
        for _ in range(3):
            c.upload(*[ic[k] for k in CONSERVED])
            c.adjoint_forward(dt, steps, 4)
            fin = c.download()
            lam0, eb = c.adjoint_backward(lamN)
            after = c.download()
            for k in CONSERVED:
                assert np.array_equal(interior(after[k]), interior(fin[k])), k
            res.append((lam0, eb))
    for lam0, eb in res[1:]:
        assert eb == res[0][1]
        for k in CONSERVED:
            assert np.array_equal(interior(lam0[k]), interior(res[0][0][k])), k


def test_kink_column_dot_product_small_step():
    """The drifting kink column over 12 steps: with a state step of 1e-8
    (eps 1e-5 on a 1e-3 direction) the central differences stay on one side
    of every limiter kink and the dot product closes to round-off; larger
    steps show the kinks (about 4e-6 at eps 1e-4)."""
    n, eta, steps = 64, 1e-3, 12
    _say(f"kink column {n}^2 dot product, {steps} steps, eps 1e-5")
    amps, phases = setups.kink_column_modes(8, 1e-3, seed=0)
    st = setups.kink_column(n, amps, phases)
    with Context(n, n, 1.0 / n, 1.0 / n, eta, cfl=0.4) as c:
        c.upload(*[st[k] for k in CONSERVED])
        dt = 0.8 * c.CFL_condition_checker()
    lamN = dict(zip(CONSERVED, _smooth_fields(n, 3)))
    v = _smooth_fields(n, 4, 1e-3)
    with Context(n, n, 1.0 / n, 1.0 / n, eta) as c:
        c.upload(*[st[k] for k in CONSERVED])
        c.adjoint_forward(dt, steps, 5)
        lam0, _ = c.adjoint_backward(lamN)

    def run(s):
        with Context(n, n, 1.0 / n, 1.0 / n, eta) as c:
            c.upload(*[s[k] for k in CONSERVED])
            c.run(1e9, dt_override=dt, max_steps=steps)
            return c.download()

    eps = 1e-5
    outs = [run({k: st[k] + sg * eps * v[q] for q, k in enumerate(CONSERVED)}) for sg in (1, -1)]
    jv = [(outs[0][k] - outs[1][k]) / (2 * eps) for k in CONSERVED]
    lhs = _dot([lam0[k] for k in CONSERVED], v)
    rhs = _dot([lamN[k] for k in CONSERVED], jv)
    assert abs(lhs - rhs) <= 1e-7 * abs(rhs), (lhs, rhs)


def test_adjoint_direct_launches_match_graph_replay():
    """With graph replay off (mhd_adjoint_use_graph(ctx, 0)), and with kernel
    event timing on, the backward pass launches its kernels directly
    (speculative solves checked afterwards) instead of replaying a graph; with
    Jacobi adjoint solves there is no CG state to check, and the result is
    the graph replay's bit for bit."""
    n, eta, steps, every = 32, 2e-2, 5, 2
    _say(f"adjoint with direct launches vs graph replay, {n}x{n}, {steps} steps")
    ic = _state(n, 21)
    lamN = dict(zip(CONSERVED, _smooth_fields(n, 22)))
    res = []
    for mode in ("graph", "timed", "direct"):
        with _ctx(ic, n, eta) as c:
            if mode == "timed":
                c.profile(["cfl"])
            if mode == "direct":
                c.adjoint_use_graph(False)
            c.adjoint_forward(4e-3, steps, every)
            res.append(c.adjoint_backward(lamN))
            if mode == "direct":     # again on the same context: still direct, still the same
                res.append(c.adjoint_backward(lamN))
    l1, e1 = res[0]
    for l2, e2 in res[1:]:
        assert e1 == e2
        for k in CONSERVED:
            assert np.array_equal(l1[k], l2[k]), k
