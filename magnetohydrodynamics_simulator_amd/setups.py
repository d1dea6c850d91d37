"""Initial conditions as padded (ny+4, nx+4) float64 arrays.

Restates the reference's problem setups (include/benchmarks.hpp of
RikibabManos/Magnetohydrodynamics_Simulator) with numpy, plus the seeded
setups of the ensemble and column workloads.  Every function returns a dict of
the six conserved fields rho, mx, my, E, bxf, byf in the reference Grid's
padded layout (ghost cells hold the setup evaluated at ghost coordinates; the
solver ignores them and refills them periodically).
"""
from __future__ import annotations

import numpy as np

PI_REF = 3.14159265359   # truncated pi of the reference (integrator_main.cpp:36, benchmarks.hpp:88,...)
GAMMA = 5.0 / 3.0
NG = 2


def grid_spacing(n: int) -> float:
    """cell_width = 2 * pi / n as set by src/integrator_main.cpp:37-38."""
    return 2 * PI_REF / n


def _from_primitive(rho, vx, vy, p, bx, by):
    """ConservedState::fromPrimitive, include/HLL.hpp:31-47."""
    v_sq = vx * vx + vy * vy
    b_sq = bx * bx + by * by
    E = p / (GAMMA - 1.0) + 0.5 * rho * v_sq + 0.5 * b_sq
    return {"rho": rho * np.ones_like(E), "mx": rho * vx, "my": rho * vy, "E": E,
            "bxf": bx * np.ones_like(E), "byf": by * np.ones_like(E)}


def _centres(nx, ny, dx, dy):
    i = np.arange(nx + 4, dtype=np.float64)
    j = np.arange(ny + 4, dtype=np.float64)
    x = (i - NG + 0.5) * dx
    y = (j - NG + 0.5) * dy
    return np.meshgrid(x, y)


def _corners(nx, ny, dx, dy, extra=1):
    i = np.arange(nx + 4 + extra, dtype=np.float64)
    j = np.arange(ny + 4 + extra, dtype=np.float64)
    return np.meshgrid((i - NG) * dx, (j - NG) * dy)


def _face_b_from_potential(Az, dx, dy):
    """bx = (Az(i,j+1) - Az(i,j))/dy, by = -(Az(i+1,j) - Az(i,j))/dx on the padded grid."""
    bx = (Az[1:, :-1] - Az[:-1, :-1]) / dy
    by = -(Az[:-1, 1:] - Az[:-1, :-1]) / dx
    return bx, by


def orszag_tang(nx: int, ny: int | None = None, dx: float | None = None, dy: float | None = None):
    """orszag_tang_initial_conditions, benchmarks.hpp:15-76: rho = 25/9,
    p = 5/3, v = (-sin y, sin x), B = curl of Az = cos(2x)/2 + cos y at corners."""
    ny = ny or nx
    dx = dx or grid_spacing(nx)
    dy = dy or grid_spacing(ny)
    XC, YC = _corners(nx, ny, dx, dy)
    Az = 0.5 * np.cos(2 * XC) + np.cos(YC)
    bx, by = _face_b_from_potential(Az, dx, dy)
    X, Y = _centres(nx, ny, dx, dy)
    return _from_primitive(25.0 / 9.0, -np.sin(Y), np.sin(X), 5.0 / 3.0, bx, by)


def decaying_sine(nx: int, ny: int | None = None, dx: float | None = None, dy: float | None = None):
    """decaying_magnetic_sine_wave_initial_conditions, benchmarks.hpp:352-417:
    Az = cos(kx x) sin(ky y) / kx at corners, rho = p = 1, v = 0."""
    ny = ny or nx
    dx = dx or grid_spacing(nx)
    dy = dy or grid_spacing(ny)
    kx = 2 * PI_REF / (nx * dx)
    ky = 2 * PI_REF / (ny * dy)
    XC, YC = _corners(nx, ny, dx, dy)
    Az = (1.0 / kx) * np.cos(kx * XC) * np.sin(ky * YC)
    bx, by = _face_b_from_potential(Az, dx, dy)
    z = np.zeros_like(bx)
    return _from_primitive(1.0, z, z, 1.0, bx, by)


def alfven(nx: int, ny: int | None = None, dx: float | None = None, dy: float | None = None):
    """alfven_wave_test_initial_values, benchmarks.hpp:78-126."""
    ny = ny or nx
    dx = dx or grid_spacing(nx)
    dy = dy or grid_spacing(ny)
    kx = 2.0 * PI_REF / (nx * dx)
    X, _ = _centres(nx, ny, dx, dy)
    by = 1e-3 * np.sin(kx * X)
    vy = -by / np.sqrt(1.0)
    return _from_primitive(1.0, np.zeros_like(X), vy, 1.0, np.ones_like(X), by)


def gaussian_pulse(nx: int, ny: int | None = None, dx: float | None = None, dy: float | None = None):
    """gaussian_density_pulse_test_initial_values, benchmarks.hpp:128-176
    (y_mid uses nx, as in the reference)."""
    ny = ny or nx
    dx = dx or grid_spacing(nx)
    dy = dy or grid_spacing(ny)
    X, Y = _centres(nx, ny, dx, dy)
    x_mid, y_mid = 0.5 * nx * dx, 0.5 * nx * dy
    d2 = (X - x_mid) * (X - x_mid) + (Y - y_mid) * (Y - y_mid)
    rho = 1.0 + np.exp(-5 * d2 / 4.0)
    one = np.ones_like(X)
    return _from_primitive(rho, one, one, 1.0, 0 * one, 0 * one)


def acoustic(nx: int, ny: int | None = None, dx: float | None = None, dy: float | None = None):
    """acoustic_wave_test_initial_values, benchmarks.hpp:178-231."""
    ny = ny or nx
    dx = dx or grid_spacing(nx)
    dy = dy or grid_spacing(ny)
    i = np.arange(nx + 4, dtype=np.float64)
    x = np.broadcast_to((i - NG) * dx, (ny + 4, nx + 4))
    W = dx * nx
    p_bg = 1.0 / GAMMA
    cs = np.sqrt(GAMMA * p_bg / 1.0)
    xs = 1e-4 * np.sin(4 * PI_REF * x / W)
    rho = 1.0 + xs + 0.0
    vx = (cs / 1.0) * xs
    p = p_bg + (cs * cs) * (xs + 0.0)
    z = np.zeros_like(rho)
    return _from_primitive(rho, vx, z, p, z, z)


def sausage(nx: int, ny: int | None = None, dx: float | None = None, dy: float | None = None,
            b0: float = 5.0, p0: float = 0.1, amp: float = 0.15):
    """sausage_instability_initial_conditions, benchmarks.hpp:233-293
    (x_mid from nxt, as in the reference)."""
    ny = ny or nx
    dx = dx or grid_spacing(nx)
    dy = dy or grid_spacing(ny)
    X, Y = _centres(nx, ny, dx, dy)
    k = 2.0 * PI_REF / (ny * dy)
    x_mid = 0.5 * (nx + 4) * dx
    r = X - x_mid
    pert = 1.0 - amp * np.cos(k * Y)
    rho = np.maximum(1.0 * np.exp(-(r * r) / (0.04 * pert)) + 0.1, 1e-4)
    p = np.maximum(p0 * np.exp(-(r * r) / (0.04 * pert)) + 0.1, 1e-4)
    by = b0 * (r / np.sqrt(0.04)) * np.exp(-(r * r) / 0.04)
    z = np.zeros_like(rho)
    return _from_primitive(rho, z, z, p, z, by)


def kink(nx: int, ny: int | None = None, dx: float | None = None, dy: float | None = None):
    """kink_instability_initial_conditions, benchmarks.hpp:295-350."""
    ny = ny or nx
    dx = dx or grid_spacing(nx)
    dy = dy or grid_spacing(ny)
    X, Y = _centres(nx, ny, dx, dy)
    r = X - 0.5 * (nx + 4) * dx
    env = np.exp(-(r * r) / 0.04)
    z = np.zeros_like(env)
    return _from_primitive(env + 0.1, 0.05 * np.sin(2.0 * PI_REF * Y) * env, z, env + 0.1, z, env)


def fourier_member_modes(rng: np.random.Generator, modes: int = 8, amp: float = 1e-2, kmax: int = 4) -> np.ndarray:
    """The seeded draws of a Fourier ensemble member: an (M, 8) table of
    kx, ky (0 < |k| <= kmax), the phases of vx, vy, Az and the amplitudes
    of vx, vy, Az (uniform in [-amp, amp]).  The table defines the member:
    fourier_member_state evaluates it on the host, the library's
    mhd_batch_init_fourier / mhd_init_fourier on the device."""
    table = np.zeros((modes, 8))
    for m in range(modes):
        while True:
            kx, ky = rng.integers(-kmax, kmax + 1, size=2)
            if 0 < kx * kx + ky * ky <= kmax * kmax:
                break
        ph = rng.uniform(0, 2 * np.pi, size=3)
        a = amp * rng.uniform(-1, 1, size=3)
        table[m] = (kx, ky, ph[0], ph[1], ph[2], a[0], a[1], a[2])
    return table


def fourier_member_state(n: int, table: np.ndarray):
    """Padded state of a Fourier ensemble member on [0, 2 pi)^2 from its mode
    table: rho = 1, p = 1, B = (1, 0.5), plus the modes on v (cell centres)
    and on B through a corner vector potential (div B = 0 to round-off)."""
    dx = grid_spacing(n)
    X, Y = _centres(n, n, dx, dx)
    XC, YC = _corners(n, n, dx, dx)
    L = n * dx
    vx = np.zeros_like(X); vy = np.zeros_like(X)
    Az = np.zeros_like(XC)
    for kx, ky, ph0, ph1, ph2, a0, a1, a2 in np.asarray(table, dtype=np.float64):
        arg = 2 * np.pi * (kx * X + ky * Y) / L
        vx += a0 * np.sin(arg + ph0)
        vy += a1 * np.sin(arg + ph1)
        k2 = (2 * np.pi / L) * np.sqrt(kx * kx + ky * ky)
        Az += (a2 / k2) * np.cos(2 * np.pi * (kx * XC + ky * YC) / L + ph2)
    Az += 1.0 * YC - 0.5 * XC           # uniform B = (1, 0.5)
    bx, by = _face_b_from_potential(Az, dx, dx)
    return _from_primitive(1.0, vx, vy, 1.0, bx, by)


def fourier_ensemble_member(n: int, rng: np.random.Generator, modes: int = 8, amp: float = 1e-2, kmax: int = 4):
    """Ensemble member on [0, 2 pi)^2: rho = 1, p = 1, B = (1, 0.5) plus seeded
    Fourier modes (|k| <= kmax, amplitude amp) on v and on B (through a vector
    potential, so div B = 0 to round-off)."""
    return fourier_member_state(n, fourier_member_modes(rng, modes, amp, kmax))


# ---------------------------------------------------------------------------
# Kink column on the unit square (the differentiation workload).  An axial
# field By = B0 G(x - 1/2) threads a dense core; density and pressure are
# displaced transversely by
#   xi(y) = sum_k a_k sin(2 pi k y + phi_k),   k = 1 .. modes,
# with seeded phases, which breaks the total-pressure balance
# p + By^2/2 = p0 + B0^2/2 of the undisplaced column.  G is the smooth
# periodic core profile exp((cos(2 pi s) - 1) / (2 pi^2 R^2)) ~ exp(-s^2/R^2).
# A weak smooth background -- flow of amplitude 0.05, a field from a corner
# potential of amplitude 0.02 (div B = 0 to round-off), 5 % density and
# pressure modulation -- keeps every reconstructed field non-uniform along
# both grid directions (with a uniform field every minmod of that field sits
# at its kink), and a uniform drift (0.3, 0.2) keeps every corner velocity
# away from zero: the upwind E_z selection of the CT step
# (solvers.hpp:326-333) jumps where a corner velocity changes sign, which
# would make the loss discontinuous in the amplitudes.  On the periodic box
# the drift only carries the column along.
# The state is analytic in the amplitudes; kink_column_dstate gives
# d(state)/d(a_k) for the gradient with respect to them.
# ---------------------------------------------------------------------------

KINK_R, KINK_B0, KINK_P0, KINK_RHO_BG = 0.2, 5.0, 0.1, 0.2
KINK_DRIFT = (0.3, 0.2)


def _kink_profile(s):
    g = np.exp((np.cos(2 * np.pi * s) - 1.0) / (2 * np.pi ** 2 * KINK_R ** 2))
    dg = -g * np.sin(2 * np.pi * s) / (np.pi * KINK_R ** 2)
    return g, dg


def kink_column_modes(modes: int = 64, amp: float = 1e-3, seed: int = 0):
    """Nominal amplitudes (all amp) and seeded phases of the displacement."""
    rng = np.random.default_rng(seed)
    return np.full(modes, amp), rng.uniform(0, 2 * np.pi, size=modes)


def _kink_xi(n, amps, phases):
    dx = 1.0 / n
    yc = (np.arange(n + 4) - NG + 0.5) * dx
    k = np.arange(1, len(amps) + 1)
    sines = np.sin(2 * np.pi * np.outer(yc, k) + phases)        # (ny+4, modes)
    return sines @ amps, sines


def _kink_background(n):
    dx = 1.0 / n
    X, Y = _centres(n, n, dx, dx)
    XC, YC = _corners(n, n, dx, dx)
    tp = 2 * np.pi
    vx = KINK_DRIFT[0] + 0.05 * np.sin(tp * Y + 0.3) + 0.03 * np.cos(2 * tp * Y + tp * X + 1.0)
    vy = KINK_DRIFT[1] + 0.05 * np.sin(tp * X + 0.7) + 0.02 * np.cos(tp * (X - Y) + 0.4)
    mod_rho = 1.0 + 0.05 * np.sin(tp * (X + 2 * Y) + 0.2)
    mod_p = 1.0 + 0.05 * np.cos(tp * (2 * X - Y) + 1.3)
    Az = 0.02 * np.sin(tp * XC + 0.5) * np.sin(tp * YC + 0.9) / tp
    bx, by = _face_b_from_potential(Az, dx, dx)
    return vx, vy, mod_rho, mod_p, bx, by


def _column_fields(n, amps, phases, sausage):
    """Fields of the displaced column.  The core profile is evaluated at
    s = x - 1/2 + xi(y) w(x): the kink shifts the core (w = -1), the sausage
    (m = 0) widens and narrows it symmetrically about x = 1/2 with the
    periodic odd weight w = -sin(2 pi (x - 1/2)) / (2 pi) ~ -(x - 1/2)."""
    dx = 1.0 / n
    xi, sines = _kink_xi(n, np.asarray(amps, dtype=np.float64), phases)
    xc = (np.arange(n + 4) - NG + 0.5) * dx
    X = xc[None, :]
    if sausage:
        w = -np.sin(2 * np.pi * (X - 0.5)) / (2 * np.pi)
        g, dg = _kink_profile(X - 0.5 + xi[:, None] * w)
    else:
        w = -np.ones_like(X)
        g, dg = _kink_profile(X - 0.5 - xi[:, None])
    gb, _ = _kink_profile(X - 0.5 + 0.0 * xi[:, None])
    vx, vy, mr, mp, bxb, byb = _kink_background(n)
    rho = (KINK_RHO_BG + g) * mr
    p = (KINK_P0 + 0.5 * KINK_B0 ** 2 * (1.0 - g * g)) * mp
    bx = bxb
    by = KINK_B0 * gb + byb
    return rho, vx, vy, p, bx, by, g, dg * w, mr, mp, sines


def _column_dstate(n, amps, phases, sausage):
    rho, vx, vy, p, bx, by, g, dg_dxi, mr, mp, sines = _column_fields(n, amps, phases, sausage)
    drho_dxi = dg_dxi * mr
    dp_dxi = -0.5 * KINK_B0 ** 2 * 2.0 * g * dg_dxi * mp
    dE_dxi = dp_dxi / (GAMMA - 1.0) + 0.5 * (vx * vx + vy * vy) * drho_dxi
    S = sines.T[:, :, None]
    return {"rho": S * drho_dxi[None], "mx": S * (vx * drho_dxi)[None], "my": S * (vy * drho_dxi)[None],
            "E": S * dE_dxi[None]}


def kink_column(n: int, amps, phases):
    """Padded state of the kink column on [0, 1]^2 (dx = dy = 1/n)."""
    rho, vx, vy, p, bx, by, *_ = _column_fields(n, amps, phases, False)
    return _from_primitive(rho, vx, vy, p, bx, by)


def kink_column_dstate(n: int, amps, phases):
    """d(state)/d(a_k) for every mode: dict of arrays (modes, ny+4, nx+4) for
    rho, mx, my and E (the face field does not depend on the amplitudes)."""
    return _column_dstate(n, amps, phases, False)


def column_gradient(lam0: dict, dstate: dict) -> np.ndarray:
    """dJ/d(a_k) = sum over cells of lam0 . d(state)/d(a_k) (physical cells)."""
    m = dstate["rho"].shape[0]
    g = np.zeros(m)
    for k in ("rho", "mx", "my", "E"):
        g += np.einsum("kij,ij->k", dstate[k][:, NG:-NG, NG:-NG], lam0[k][NG:-NG, NG:-NG])
    return g


kink_column_gradient = column_gradient


# ---------------------------------------------------------------------------
# Sausage column on the unit square (the batched differentiation workload):
# the kink column's core, field, background and drift, with the core widened
# and narrowed symmetrically about x = 1/2 (m = 0) by xi(y), seeded modes of
# total amplitude 1e-2.  Ensembles draw the side from {256, 512, 768, 1024}
# and eta log-uniformly from [1e-4, 1e-2].
# ---------------------------------------------------------------------------

SAUSAGE_SIDES = (256, 512, 768, 1024)


def sausage_column_modes(modes: int = 8, amp: float = 1e-2, seed: int = 0):
    """Amplitudes amp / modes each (|xi| <= amp) and seeded phases."""
    rng = np.random.default_rng(seed)
    return np.full(modes, amp / modes), rng.uniform(0, 2 * np.pi, size=modes)


def sausage_column(n: int, amps, phases):
    """Padded state of the sausage column on [0, 1]^2 (dx = dy = 1/n)."""
    rho, vx, vy, p, bx, by, *_ = _column_fields(n, amps, phases, True)
    return _from_primitive(rho, vx, vy, p, bx, by)


def sausage_column_dstate(n: int, amps, phases):
    """d(state)/d(a_k) of the sausage column (as kink_column_dstate)."""
    return _column_dstate(n, amps, phases, True)


def sausage_ensemble(nruns: int = 64, seed: int = 0, sides=SAUSAGE_SIDES, eta_range=(1e-4, 1e-2)):
    """Seeded ensemble: side from `sides`, eta log-uniform in eta_range and a
    phase seed per run.  Returns a list of (n, eta, phase_seed); the state is
    sausage_column(n, *sausage_column_modes(seed=phase_seed))."""
    rng = np.random.default_rng(seed)
    ns = rng.choice(np.asarray(sides), size=nruns)
    etas = 10.0 ** rng.uniform(np.log10(eta_range[0]), np.log10(eta_range[1]), size=nruns)
    seeds = rng.integers(0, 2**31 - 1, size=nruns)
    return [(int(n), float(e), int(s)) for n, e, s in zip(ns, etas, seeds)]


def window_magnetic_energy(state: dict, n: int, dx: float, lo: float = 0.25, hi: float = 0.75):
    """Loss of the differentiation workload: (1/2) sum (bxf^2 + byf^2) dx dy
    over the faces of the central window [lo, hi)^2, and its gradient
    with respect to the padded state."""
    i0, i1 = int(round(lo * n)), int(round(hi * n))
    grad = {k: np.zeros_like(state["rho"]) for k in ("rho", "mx", "my", "E", "bxf", "byf")}
    loss = 0.0
    for k in ("bxf", "byf"):
        w = state[k][NG + i0:NG + i1, NG + i0:NG + i1]
        loss += 0.5 * float(np.sum(w * w)) * dx * dx
        grad[k][NG + i0:NG + i1, NG + i0:NG + i1] = w * dx * dx
    return loss, grad


def window_magnetic_energy_difference(plus: dict, minus: dict, n: int, dx: float, lo: float = 0.25,
                                      hi: float = 0.75) -> float:
    """window_magnetic_energy(plus) - window_magnetic_energy(minus) summed as
    (1/2) sum (B+ - B-)(B+ + B-) dx dy over the same faces: the same number
    without the cancellation of two O(1) sums (central differences of the
    loss then carry the rounding of the two runs, not that of the loss)."""
    i0, i1 = int(round(lo * n)), int(round(hi * n))
    d = 0.0
    for k in ("bxf", "byf"):
        a = plus[k][NG + i0:NG + i1, NG + i0:NG + i1]
        b = minus[k][NG + i0:NG + i1, NG + i0:NG + i1]
        d += 0.5 * float(np.sum((a - b) * (a + b))) * dx * dx
    return d


ENSEMBLE_SIDES = (64, 96, 128, 192, 256, 384, 512)


def ensemble(nruns: int = 256, seed: int = 0, sides=ENSEMBLE_SIDES, eta_range=(1e-4, 1e-1)):
    """Seeded ensemble of independent periodic runs: side drawn from `sides`,
    eta log-uniform in eta_range, initial state fourier_ensemble_member.
    Returns a list of (n, eta, rng_seed_of_member); build the state with
    ensemble_state(n, member_seed)."""
    rng = np.random.default_rng(seed)
    ns = rng.choice(np.asarray(sides), size=nruns)
    etas = 10.0 ** rng.uniform(np.log10(eta_range[0]), np.log10(eta_range[1]), size=nruns)
    member_seeds = rng.integers(0, 2**31 - 1, size=nruns)
    return [(int(n), float(e), int(ms)) for n, e, ms in zip(ns, etas, member_seeds)]


def ensemble_state(n: int, member_seed: int, **kw):
    return fourier_ensemble_member(n, np.random.default_rng(member_seed), **kw)


def ensemble_modes(member_seed: int, **kw) -> np.ndarray:
    """Mode table of ensemble member `member_seed` (for the device generator)."""
    return fourier_member_modes(np.random.default_rng(member_seed), **kw)


def shock_tube(nx: int, ny: int = 4, dx: float | None = None, dy: float | None = None, x0: float = 0.5):
    """Brio & Wu's MHD shock tube along x on [0, 1) (dx = dy = 1 / nx unless
    given): rho = 1, p = 1, by = 1 left of x0; rho = 0.125, p = 0.1, by = -1
    right of it; bx = 0.75, v = 0.  The solver's gamma is 5/3 (Brio & Wu use
    2), so the wave pattern is that of the same Riemann problem at 5/3.  Run
    with outflow boundaries in x (the reference cannot: it is periodic only,
    Grid.hpp:141-142)."""
    dx = dx if dx is not None else 1.0 / nx
    dy = dy if dy is not None else dx
    X, _ = _centres(nx, ny, dx, dy)
    left = X < x0
    rho = np.where(left, 1.0, 0.125)
    p = np.where(left, 1.0, 0.1)
    by = np.where(left, 1.0, -1.0)
    bx = np.full_like(X, 0.75)
    zero = np.zeros_like(X)
    return _from_primitive(rho, zero, zero, p, bx, by)


SETUPS = {
    "orszag_tang": orszag_tang,
    "decaying_sine": decaying_sine,
    "alfven": alfven,
    "gaussian": gaussian_pulse,
    "acoustic": acoustic,
    "sausage": sausage,
    "kink": kink,
}
