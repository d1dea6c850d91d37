// mhd_device.hpp -- per-cell / per-face FP64 device math of the MHD step.
//
// These are the scalar building blocks that the fused stage kernel strings
// together.  Each follows a reference function (paths relative to the
// reference repository) and keeps its operation order, except where the
// reference's own -ffast-math build (CMakeLists.txt:21) already rewrites a
// division: GCC's -freciprocal-math turns x / c for a constant c into x * (1/c)
// and replaces three or more divisions by the same value with one reciprocal
// and multiplies; the same rewrites are made here, so both sides round alike
// to within an ulp per operation.
#pragma once

#include <hip/hip_runtime.h>

namespace mhd {

constexpr double kGamma = 5.0 / 3.0;                    // Grid.hpp:104, HLL.hpp:31
constexpr double kGammaM1 = kGamma - 1.0;                // (gamma - 1)
constexpr double kInvGammaM1 = 1.0 / (kGamma - 1.0);     // x / (gamma - 1.0) -> x * (1 / (gamma - 1))
constexpr double kFloor = 1e-8;                          // Grid.hpp:115,124
constexpr double kTie = 1e-12;                           // solvers.hpp:326,333

// Primitive state, PrimitiveState (HLL.hpp:12-19).
struct Prim {
  double rho, vx, vy, p, bx, by;
};

// Flux through one face: mass, x/y momentum, energy and the one non-zero
// induction component (F_by on x-faces, G_bx on y-faces); the normal
// induction flux is identically zero (HLL.hpp:76,99,166,212).
struct Flux {
  double rho, mx, my, E, b;
};

// minmod, MUSCL.hpp:11-17.
__device__ __forceinline__ double minmod(double a, double b) {
  return (a * b > 0.0) ? ((fabs(a) < fabs(b)) ? a : b) : 0.0;
}

// MUSCLReconstructOne, MUSCL.hpp:26-47: undivided minmod slope, L = low face, R = high face.
__device__ __forceinline__ void muscl(double fm, double f0, double fp, double& L, double& R) {
  double s = minmod(f0 - fm, fp - f0);
  L = f0 - 0.5 * s;
  R = f0 + 0.5 * s;
}

// std::min / std::max semantics.
__device__ __forceinline__ double dmin(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double dmax(double a, double b) { return (a < b) ? b : a; }

// 1/x for finite, normal, non-zero x: the hardware estimate (v_rcp_f64)
// refined by two Newton-Raphson steps r <- r + r (1 - x r), each of which
// doubles the number of correct bits; within about an ulp of the correctly
// rounded reciprocal.  The kernels apply it to densities (floored at 1e-8)
// and to HLL wave-speed spans (> 0 where used), so the scaling and special-
// case handling of the IEEE division sequence are not needed.
__device__ __forceinline__ double rcp(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}

// sqrt(x) for normal x > 0: Goldschmidt's coupled iteration on
// g ~ sqrt(x), h ~ 1 / (2 sqrt(x)) started from the hardware reciprocal
// square root (v_rsq_f64), followed by two Newton corrections
// g <- g + h (x - g^2).
__device__ __forceinline__ double sqrt_nr(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y, h = 0.5 * y;
  const double r = fma(-h, g, 0.5);
  g = fma(g, r, g);
  h = fma(h, r, h);
  double d = fma(-g, g, x);
  g = fma(d, h, g);
  d = fma(-g, g, x);
  return fma(d, h, g);
}

// Unfloored pressure of Grid::computePrimitives (Grid.hpp:121-126).  FMA
// contraction is off here, so every kernel that evaluates it -- the explicit
// stage and the adjoint kernels that recompute the forward's branch choices
// -- rounds alike whatever code it is inlined into (a uniform pressure is
// recovered with rounding noise, and the minmod branch of its slopes follows
// that noise).
__device__ __forceinline__ double pressure_raw(double E, double rho_f, double vx, double vy, double bxc,
                                               double byc) {
#pragma clang fp contract(off)
  const double kinE = 0.5 * rho_f * (vx * vx + vy * vy);
  const double magE = (bxc * bxc + byc * byc) / 2;
  const double intE = E - kinE - magE;
  return kGammaM1 * intE;
}

// Grid::computePrimitives for one cell, Grid.hpp:111-126.  rho is the stored
// density; the velocities use it before the floor is applied.
__device__ __forceinline__ void primitives(double rho, double mx, double my, double E, double bxc, double byc,
                                           double& rho_f, double& vx, double& vy, double& p) {
  const double inv = rcp(rho);
  vx = mx * inv;
  vy = my * inv;
  rho_f = (rho < kFloor) ? kFloor : rho;
  p = pressure_raw(E, rho_f, vx, vy, bxc, byc);
  if (p < kFloor) p = kFloor;
}

// getFastWaveSpeed, HLL.hpp:107-122 (three divisions by rho -> one reciprocal).
__device__ __forceinline__ double fast_speed(const Prim& s, double bn) {
  double B_sq = s.bx * s.bx + s.by * s.by;
  double inv_rho = rcp(s.rho);
  double sound_v_sq = (kGamma * s.p) * inv_rho;
  double tot_alfven_v_sq = B_sq * inv_rho;
  double norm_alfven_v_sq = (bn * bn) * inv_rho;
  double sum = tot_alfven_v_sq + sound_v_sq;
  double term = sum * sum - 4.0 * sound_v_sq * norm_alfven_v_sq;
  // term >= 0; a term below 2^-1000 adds nothing to sum (sum >= gamma p / rho
  // with both floors applied), so its root is taken as 0
  const double root = (term >= 0x1p-1000) ? sqrt_nr(term) : 0.0;
  return sqrt_nr(0.5 * (sum + root));
}

// convertPrimitiveX (HLL.hpp:61-81) and the conserved energy of
// ConservedState::fromPrimitive (HLL.hpp:31-47) for one state.
__device__ __forceinline__ void phys_flux_x(const Prim& s, Flux& F, double& E_cons) {
  double v_sq = s.vx * s.vx + s.vy * s.vy;
  double B_sq = s.bx * s.bx + s.by * s.by;
  double P_tot = s.p + 0.5 * B_sq;
  double v_dot_B = s.vx * s.bx + s.vy * s.by;
  double pe = s.p * kInvGammaM1;
  double E = pe + 0.5 * (s.rho * v_sq + B_sq);
  F.rho = s.rho * s.vx;
  F.mx = s.rho * s.vx * s.vx + P_tot - s.bx * s.bx;
  F.my = s.rho * s.vx * s.vy - s.bx * s.by;
  F.E = (E + P_tot) * s.vx - s.bx * v_dot_B;
  F.b = s.vx * s.by - s.vy * s.bx;
  E_cons = pe + 0.5 * s.rho * v_sq + 0.5 * B_sq;
}

// convertPrimitiveY (HLL.hpp:84-104) and fromPrimitive's energy.
__device__ __forceinline__ void phys_flux_y(const Prim& s, Flux& F, double& E_cons) {
  double v_sq = s.vx * s.vx + s.vy * s.vy;
  double B_sq = s.bx * s.bx + s.by * s.by;
  double P_tot = s.p + 0.5 * B_sq;
  double v_dot_B = s.vx * s.bx + s.vy * s.by;
  double pe = s.p * kInvGammaM1;
  double E = pe + 0.5 * (s.rho * v_sq + B_sq);
  F.rho = s.rho * s.vy;
  F.my = s.rho * s.vy * s.vy + P_tot - s.by * s.by;
  F.mx = s.rho * s.vx * s.vy - s.bx * s.by;
  F.E = (E + P_tot) * s.vy - s.by * v_dot_B;
  F.b = s.vy * s.bx - s.vx * s.by;
  E_cons = pe + 0.5 * s.rho * v_sq + 0.5 * B_sq;
}

// computeHLLFluxX, HLL.hpp:126-172.  L/R are the states left/right of the
// face; both carry the same face-centred normal field in .bx.
__device__ __forceinline__ Flux hll_x(const Prim& L, const Prim& R) {
  Flux FL, FR;
  double EL, ER;
  phys_flux_x(L, FL, EL);
  phys_flux_x(R, FR, ER);
  double cL = fast_speed(L, L.bx);
  double cR = fast_speed(R, R.bx);
  double sL = dmin(L.vx - cL, R.vx - cR);
  double sR = dmax(L.vx + cL, R.vx + cR);
  if (sL >= 0) return FL;
  if (sR <= 0) return FR;
  double inv_S = rcp(sR - sL);
  double sLR = sL * sR;
  Flux H;
  H.rho = (sR * FL.rho - sL * FR.rho + sLR * (R.rho - L.rho)) * inv_S;
  H.mx = (sR * FL.mx - sL * FR.mx + sLR * (R.rho * R.vx - L.rho * L.vx)) * inv_S;
  H.my = (sR * FL.my - sL * FR.my + sLR * (R.rho * R.vy - L.rho * L.vy)) * inv_S;
  H.E = (sR * FL.E - sL * FR.E + sLR * (ER - EL)) * inv_S;
  H.b = (sR * FL.b - sL * FR.b + sLR * (R.by - L.by)) * inv_S;
  return H;
}

// computeHLLFluxY, HLL.hpp:177-218.  D/U are the states below/above the face;
// both carry the same face-centred normal field in .by.
__device__ __forceinline__ Flux hll_y(const Prim& D, const Prim& U) {
  Flux FD, FU;
  double ED, EU;
  phys_flux_y(D, FD, ED);
  phys_flux_y(U, FU, EU);
  double cD = fast_speed(D, D.by);
  double cU = fast_speed(U, U.by);
  double sD = dmin(D.vy - cD, U.vy - cU);
  double sU = dmax(D.vy + cD, U.vy + cU);
  if (sD >= 0) return FD;
  if (sU <= 0) return FU;
  double inv_S = rcp(sU - sD);
  double sDU = sD * sU;
  Flux H;
  H.rho = (sU * FD.rho - sD * FU.rho + sDU * (U.rho - D.rho)) * inv_S;
  H.mx = (sU * FD.mx - sD * FU.mx + sDU * (U.rho * U.vx - D.rho * D.vx)) * inv_S;
  H.my = (sU * FD.my - sD * FU.my + sDU * (U.rho * U.vy - D.rho * D.vy)) * inv_S;
  H.E = (sU * FD.E - sD * FU.E + sDU * (EU - ED)) * inv_S;
  H.b = (sU * FD.b - sD * FU.b + sDU * (U.bx - D.bx)) * inv_S;
  return H;
}

// Upwind corner E_z, cell_corner_electric_fields (solvers.hpp:314-340).
// eN/eS/eE/eW are the four face electric fields around the corner, vxc/vyc
// the four-cell average velocity.  The reference's abs() is a double abs in
// its real build (Eigen includes the libstdc++ <stdlib.h> wrapper), hence fabs.
__device__ __forceinline__ double corner_ez(double eN, double eS, double eE, double eW, double vxc, double vyc) {
  double ezy = (fabs(vyc) < kTie) ? 0.5 * (eS + eN) : ((vyc > 0.0) ? eS : eN);
  double ezx = (fabs(vxc) < kTie) ? 0.5 * (eE + eW) : ((vxc > 0.0) ? eW : eE);
  return 0.5 * (ezx + ezy);
}

// Virtual block index: consecutive workgroup ids are dealt to the 8 XCDs in
// turn; remapping gives each XCD (and its L2) a contiguous run of tiles.
__device__ __forceinline__ int xcd_block(int b, int nb) {
  if ((nb & 7) != 0) return b;
  return (b & 7) * (nb >> 3) + (b >> 3);
}

// Periodic index helpers.
__device__ __forceinline__ int pmod(int a, int n) {
  int r = a % n;
  return r < 0 ? r + n : r;
}
__device__ __forceinline__ int wrap1(int a, int n) { return a < 0 ? a + n : (a >= n ? a - n : a); }

// Index k on an axis of n points.  A periodic axis wraps (any k for bc_mod,
// |k| < 2n for bc_wrap); an outflow axis takes the nearest index inside the
// domain: every field is extended by its boundary value (the zero-gradient
// ghost rule), so a corner unknown beyond the last corner equals the last.
__device__ __forceinline__ int bc_wrap(int k, int n, int open) {
  return open ? (k < 0 ? 0 : (k >= n ? n - 1 : k)) : wrap1(k, n);
}
__device__ __forceinline__ int bc_mod(int k, int n, int open) {
  return open ? (k < 0 ? 0 : (k >= n ? n - 1 : k)) : pmod(k, n);
}

}  // namespace mhd
