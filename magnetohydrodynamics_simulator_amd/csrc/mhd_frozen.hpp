// mhd_frozen.hpp -- the explicit right-hand side with its branch choices held
// fixed, for checking the adjoint against finite differences.
//
// The explicit right-hand side L (explicit_rhs, solvers.hpp:393-402) is
// piecewise smooth: minmod (MUSCL.hpp:11-17) picks one of its two slopes or
// zero, HLL (HLL.hpp:145-152) picks F_L, F_R or the averaged flux and takes a
// min and a max of wave speeds, the corner E_z (solvers.hpp:326-338) picks an
// upwind face or the average, and the density and pressure floors
// (Grid.hpp:115,124) switch between a value and a constant.  The adjoint
// differentiates L with each choice taken as at the recorded forward state,
// so it is the derivative of the map in which those choices are frozen.  This
// header holds the two halves of that map:
//   * record_cell: the choices at one cell of a state, encoded in 64 bits,
//     from the same device functions and inputs the adjoint's gather kernels
//     use (so the record is the adjoint's own linearisation);
//   * fz_prims_cell / fz_faces_cell / fz_update_cell: L with every choice
//     read from such a record (the frozen map; a smooth function of the
//     state, equal to L wherever the state makes the same choices).
//
// Branch word of cell (i, j) at one stage:
//   bits  0..19  minmod choice of the 10 slopes of the cell (2 bits each:
//                0 zero, 1 the backward difference a, 2 the forward one b);
//                slope q = 0..4: x-reconstruction of rho, vx, vy, p, byc;
//                q = 5..9: y-reconstruction of rho, vx, vy, p, bxc
//   bits 20..25  HLL of the cell's right x-face, bits 26..31 of its top
//                y-face (y-faces as x-faces of the x<->y swapped states):
//                bits 0..1 case (0 averaged, 1 F_L, 2 F_R), bit 2 S_L taken
//                from the right state, bit 3 S_R from the right state,
//                bits 4, 5 the fast-speed inner root of L, R is non-zero
//   bits 32..35  corner E_z of the cell's lower-left corner: bits 0..1 the
//                y choice (0 average, 1 south, 2 north), bits 2..3 the x
//                choice (0 average, 1 west, 2 east)
//   bit  36      density floor of the primitives (stored rho < 1e-8)
//   bit  37      pressure floor (unfloored pressure < 1e-8)
//   bit  38      density floor of the state the stage produces (S* after
//                stage 1, U3* after stage 2)
#pragma once

#include <cstdint>

#include "mhd_adjoint_lin.hpp"
#include "mhd_adjoint_math.hpp"
#include "mhd_device.hpp"

namespace mhd {

__device__ __forceinline__ uint64_t mm_code(double a, double b) {
  return (a * b > 0.0) ? ((fabs(a) < fabs(b)) ? 1u : 2u) : 0u;
}

// Non-zero inner root of getFastWaveSpeed as fast_speed_vjp decides it.
__device__ __forceinline__ uint64_t root_code(const Prim& s, double bn) {
  const double B2 = s.bx * s.bx + s.by * s.by;
  const double ir = 1.0 / s.rho;
  const double a2 = (kGamma * s.p) * ir;
  const double b2 = B2 * ir;
  const double n2 = (bn * bn) * ir;
  const double sum = b2 + a2;
  const double term = sum * sum - 4.0 * a2 * n2;
  return (term >= 0x1p-1000) ? 1u : 0u;
}

// HLL x choices as hll_x_vjp takes them.
__device__ __forceinline__ uint64_t hll_code(const Prim& L, const Prim& R) {
  const double cL = fast_speed(L, L.bx);
  const double cR = fast_speed(R, R.bx);
  const double aL = L.vx - cL, aR = R.vx - cR;
  const double bL = L.vx + cL, bR = R.vx + cR;
  const double sL = dmin(aL, aR);
  const double sR = dmax(bL, bR);
  uint64_t c = (sL >= 0) ? 1u : ((sR <= 0) ? 2u : 0u);
  if (aR < aL) c |= 4u;
  if (bL < bR) c |= 8u;
  c |= root_code(L, L.bx) << 4;
  c |= root_code(R, R.bx) << 5;
  return c;
}

// Records the branch word of cell (i, j) of the state a.s, whose primitives
// ka_prims has written to a.P; rho_after: the density of the state the stage
// produces.
__device__ __forceinline__ uint64_t record_cell(const AdjArgs& a, const double* rho_after, int i, int j) {
  const int nx = a.nx, ny = a.ny;
  const int o = j * nx + i;
  const int ip = wp(i + 1, nx), jp = wp(j + 1, ny);
  uint64_t w = 0;
  // minmod choices
  for (int dir = 0; dir < 2; ++dir) {
    const int om = (dir == 0) ? j * nx + wp(i - 1, nx) : wp(j - 1, ny) * nx + i;
    const int op = (dir == 0) ? j * nx + ip : jp * nx + i;
    const int tb = (dir == 0) ? 5 : 4;
    const int vars[5] = {0, 1, 2, 3, tb};
#pragma unroll
    for (int q = 0; q < 5; ++q) {
      const double* f = a.P[vars[q]];
      w |= mm_code(f[o] - f[om], f[op] - f[o]) << (2 * (5 * dir + q));
    }
  }
  // HLL of the right x-face and the top y-face (the states of ka_faces_cell)
  {
    double lo[5], hi[5], lo2[5], hi2[5];
    recon(a, 0, i, j, lo, hi);
    recon(a, 0, ip, j, lo2, hi2);
    const double bn = a.s[4][j * nx + ip];
    const Prim L{hi[0], hi[1], hi[2], hi[3], bn, hi[4]};
    const Prim R{lo2[0], lo2[1], lo2[2], lo2[3], bn, lo2[4]};
    w |= hll_code(L, R) << 20;
  }
  {
    double lo[5], hi[5], lo2[5], hi2[5];
    recon(a, 1, i, j, lo, hi);
    recon(a, 1, i, jp, lo2, hi2);
    const double bn = a.s[5][jp * nx + i];
    const Prim D{hi[0], hi[1], hi[2], hi[3], hi[4], bn};
    const Prim U{lo2[0], lo2[1], lo2[2], lo2[3], lo2[4], bn};
    w |= hll_code(swap_xy(D), swap_xy(U)) << 26;
  }
  // corner E_z weights of corner (i, j)
  double wS, wN, wW, wE;
  ez_weights(a, i, j, wS, wN, wW, wE);
  const uint64_t cy = (wS == 0.5) ? 0u : (wS == 1.0 ? 1u : 2u);
  const uint64_t cx = (wW == 0.5) ? 0u : (wW == 1.0 ? 1u : 2u);
  w |= (cy | (cx << 2)) << 32;
  // floors
  if (a.s[0][o] < kFloor) w |= 1ull << 36;
  if (a.P[6][o] == 0.0) w |= 1ull << 37;
  if (!(rho_after[o] > kFloor)) w |= 1ull << 38;
  return w;
}

// ---------------------------------------------------------------------------
// The frozen right-hand side.
// ---------------------------------------------------------------------------

struct FrozenArgs {
  int nx, ny;
  double idx, idy;
  const double* s[6];     // state
  const uint64_t* br;     // branch words of this stage
  double* P[6];           // primitives rho_f, vx, vy, p, bxc, byc
  double* fx[5];          // flux through the right x-face of each cell
  double* fy[5];          // flux through the top y-face
  double* out[6];         // the right-hand side
};

__device__ __forceinline__ void fz_prims_cell(const FrozenArgs& a, int i, int j) {
  const int nx = a.nx, ny = a.ny;
  const int o = j * nx + i;
  const uint64_t w = a.br[o];
  const double bxc = (a.s[4][o] + a.s[4][j * nx + wp(i + 1, nx)]) / 2.0;
  const double byc = (a.s[5][o] + a.s[5][wp(j + 1, ny) * nx + i]) / 2.0;
  const double rho = a.s[0][o];
  const double inv = rcp(rho);
  const double vx = a.s[1][o] * inv, vy = a.s[2][o] * inv;
  const double rf = ((w >> 36) & 1u) ? kFloor : rho;
  const double praw = pressure_raw(a.s[3][o], rf, vx, vy, bxc, byc);
  a.P[0][o] = rf;
  a.P[1][o] = vx;
  a.P[2][o] = vy;
  a.P[3][o] = ((w >> 37) & 1u) ? kFloor : praw;
  a.P[4][o] = bxc;
  a.P[5][o] = byc;
}

// MUSCLReconstructOne with the slope choice given.
__device__ __forceinline__ void muscl_frozen(double fm, double f0, double fp, uint64_t code, double& L, double& R) {
  const double s = (code == 1u) ? (f0 - fm) : ((code == 2u) ? (fp - f0) : 0.0);
  L = f0 - 0.5 * s;
  R = f0 + 0.5 * s;
}

// The x (dir 0) or y (dir 1) reconstruction of cell (i, j) with its recorded slopes.
__device__ __forceinline__ void recon_frozen(const FrozenArgs& a, int dir, int i, int j, double (&lo)[5],
                                             double (&hi)[5]) {
  const int nx = a.nx, ny = a.ny;
  int om, o0, op;
  if (dir == 0) { om = j * nx + wp(i - 1, nx); o0 = j * nx + i; op = j * nx + wp(i + 1, nx); }
  else { om = wp(j - 1, ny) * nx + i; o0 = j * nx + i; op = wp(j + 1, ny) * nx + i; }
  const uint64_t w = a.br[o0];
  const int tb = (dir == 0) ? 5 : 4;
  const int vars[5] = {0, 1, 2, 3, tb};
#pragma unroll
  for (int q = 0; q < 5; ++q) {
    const double* f = a.P[vars[q]];
    muscl_frozen(f[om], f[o0], f[op], (w >> (2 * (5 * dir + q))) & 3u, lo[q], hi[q]);
  }
}

// getFastWaveSpeed with the inner root's branch given (a root that the
// perturbed state would make imaginary is taken as 0).
__device__ __forceinline__ double fast_speed_frozen(const Prim& s, double bn, uint64_t root) {
  const double B_sq = s.bx * s.bx + s.by * s.by;
  const double inv_rho = rcp(s.rho);
  const double sound_v_sq = (kGamma * s.p) * inv_rho;
  const double tot_alfven_v_sq = B_sq * inv_rho;
  const double norm_alfven_v_sq = (bn * bn) * inv_rho;
  const double sum = tot_alfven_v_sq + sound_v_sq;
  const double term = sum * sum - 4.0 * sound_v_sq * norm_alfven_v_sq;
  const double r = root ? sqrt(fmax(term, 0.0)) : 0.0;
  return sqrt(0.5 * (sum + r));
}

// computeHLLFluxX with its case and wave-speed choices given.
__device__ __forceinline__ Flux hll_x_frozen(const Prim& L, const Prim& R, uint64_t code) {
  Flux FL, FR;
  double EL, ER;
  phys_flux_x(L, FL, EL);
  phys_flux_x(R, FR, ER);
  const uint64_t cs = code & 3u;
  if (cs == 1u) return FL;
  if (cs == 2u) return FR;
  const double cL = fast_speed_frozen(L, L.bx, (code >> 4) & 1u);
  const double cR = fast_speed_frozen(R, R.bx, (code >> 5) & 1u);
  const double sL = (code & 4u) ? R.vx - cR : L.vx - cL;
  const double sR = (code & 8u) ? R.vx + cR : L.vx + cL;
  const double inv_S = 1.0 / (sR - sL);
  const double sLR = sL * sR;
  Flux H;
  H.rho = (sR * FL.rho - sL * FR.rho + sLR * (R.rho - L.rho)) * inv_S;
  H.mx = (sR * FL.mx - sL * FR.mx + sLR * (R.rho * R.vx - L.rho * L.vx)) * inv_S;
  H.my = (sR * FL.my - sL * FR.my + sLR * (R.rho * R.vy - L.rho * L.vy)) * inv_S;
  H.E = (sR * FL.E - sL * FR.E + sLR * (ER - EL)) * inv_S;
  H.b = (sR * FL.b - sL * FR.b + sLR * (R.by - L.by)) * inv_S;
  return H;
}

// Fluxes of the right x-face and the top y-face of cell (i, j).
__device__ __forceinline__ void fz_faces_cell(const FrozenArgs& a, int i, int j) {
  const int nx = a.nx, ny = a.ny;
  const int o = j * nx + i;
  const int ip = wp(i + 1, nx), jp = wp(j + 1, ny);
  const uint64_t w = a.br[o];
  {
    double lo[5], hi[5], lo2[5], hi2[5];
    recon_frozen(a, 0, i, j, lo, hi);
    recon_frozen(a, 0, ip, j, lo2, hi2);
    const double bn = a.s[4][j * nx + ip];
    const Prim L{hi[0], hi[1], hi[2], hi[3], bn, hi[4]};
    const Prim R{lo2[0], lo2[1], lo2[2], lo2[3], bn, lo2[4]};
    const Flux F = hll_x_frozen(L, R, (w >> 20) & 63u);
    a.fx[0][o] = F.rho; a.fx[1][o] = F.mx; a.fx[2][o] = F.my; a.fx[3][o] = F.E; a.fx[4][o] = F.b;
  }
  {
    double lo[5], hi[5], lo2[5], hi2[5];
    recon_frozen(a, 1, i, j, lo, hi);
    recon_frozen(a, 1, i, jp, lo2, hi2);
    const double bn = a.s[5][jp * nx + i];
    const Prim D{hi[0], hi[1], hi[2], hi[3], hi[4], bn};
    const Prim U{lo2[0], lo2[1], lo2[2], lo2[3], lo2[4], bn};
    // the y-flux is the x-flux of the swapped states with mx and my exchanged
    const Flux G = hll_x_frozen(swap_xy(D), swap_xy(U), (w >> 26) & 63u);
    a.fy[0][o] = G.rho; a.fy[1][o] = G.my; a.fy[2][o] = G.mx; a.fy[3][o] = G.E; a.fy[4][o] = G.b;
  }
}

// Corner E_z of corner (i, j) with its upwind choice given.
__device__ __forceinline__ double fz_corner(const FrozenArgs& a, int i, int j) {
  const int nx = a.nx, ny = a.ny;
  const int im = wp(i - 1, nx), jm = wp(j - 1, ny);
  const uint64_t c = (a.br[j * nx + i] >> 32) & 15u;
  const double eN = -a.fx[4][j * nx + im];
  const double eS = -a.fx[4][jm * nx + im];
  const double eE = a.fy[4][jm * nx + i];
  const double eW = a.fy[4][jm * nx + im];
  const uint64_t cy = c & 3u, cx = c >> 2;
  const double ezy = (cy == 0u) ? 0.5 * (eS + eN) : (cy == 1u ? eS : eN);
  const double ezx = (cx == 0u) ? 0.5 * (eE + eW) : (cx == 1u ? eW : eE);
  return 0.5 * (ezx + ezy);
}

// time_derivatives (solvers.hpp:253-292) of cell (i, j).
__device__ __forceinline__ void fz_update_cell(const FrozenArgs& a, int i, int j) {
  const int nx = a.nx, ny = a.ny;
  const int o = j * nx + i;
  const int im = j * nx + wp(i - 1, nx), jm = wp(j - 1, ny) * nx + i;
#pragma unroll
  for (int q = 0; q < 4; ++q)
    a.out[q][o] = -((a.fx[q][o] - a.fx[q][im]) * a.idx + (a.fy[q][o] - a.fy[q][jm]) * a.idy);
  const double ez = fz_corner(a, i, j);
  a.out[4][o] = -(fz_corner(a, i, wp(j + 1, ny)) - ez) * a.idy;
  a.out[5][o] = (fz_corner(a, wp(i + 1, nx), j) - ez) * a.idx;
}

// Stage 1 combinations: S* = U + c1 f (density floor as recorded), W = U + cw f.
__device__ __forceinline__ void fz_comb1_cell(int o, const double* const* u, const double* const* f, double c1,
                                              double cw, const uint64_t* br, double* const* s, double* const* w) {
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    double v = u[q][o] + c1 * f[q][o];
    if (q == 0 && ((br[o] >> 38) & 1u)) v = kFloor;
    s[q][o] = v;
    w[q][o] = u[q][o] + cw * f[q][o];
  }
}

// Stage 2 combination: U3* = W + c2 f + c3 R(q) on B (density floor as recorded).
__device__ __forceinline__ void fz_comb2_cell(const LinArgs& g, int i, int j, const double* const* w,
                                              const double* const* f, const double* q, double c2, double c3,
                                              const uint64_t* br, double* const* u) {
  const int o = j * g.nx + i;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    double v = w[k][o] + c2 * f[k][o];
    if (k == 0 && ((br[o] >> 38) & 1u)) v = kFloor;
    u[k][o] = v;
  }
  const double qc = q[o], qn = q[wp(j + 1, g.ny) * g.nx + i], qe = q[j * g.nx + wp(i + 1, g.nx)];
  u[4][o] = w[4][o] + c2 * f[4][o] + c3 * (-(qn - qc) * g.idy);
  u[5][o] = w[5][o] + c2 * f[5][o] + c3 * ((qe - qc) * g.idx);
}

}  // namespace mhd
