// mhd_adjoint_math.hpp -- reverse-mode (vector-Jacobian) products of the
// per-face device math of mhd_device.hpp (physical flux, fast wave speed, HLL
// x/y, MUSCL-minmod) and the per-cell bodies of the explicit right-hand
// side's adjoint built from them (run by the single-run kernels of
// mhd_adjoint.hip and the batched ones of mhd_batch_adjoint.hip).
#pragma once

#include "mhd_device.hpp"

namespace mhd {

// ---------------------------------------------------------------------------
// Reverse-mode building blocks.  Each *_vjp recomputes the forward values it
// needs (with the forward kernels' device functions, so branch decisions match
// the forward pass) and accumulates the input adjoints for an output adjoint.
// ---------------------------------------------------------------------------

// phys_flux_x (HLL.hpp:61-81) and the conserved energy (HLL.hpp:31-47):
// F = (rho vx, rho vx^2 + Pt - bx^2, rho vx vy - bx by, (E + Pt) vx - bx v.B,
// vx by - vy bx), Ec = p/(g-1) + rho v^2/2 + B^2/2.  Adjoint of (F, Ec).
__device__ __forceinline__ void phys_flux_x_vjp(const Prim& s, const double Fb[5], double Eb, double (&sb)[6]) {
  const double v2 = s.vx * s.vx + s.vy * s.vy;
  const double B2 = s.bx * s.bx + s.by * s.by;
  const double Pt = s.p + 0.5 * B2;
  const double vB = s.vx * s.bx + s.vy * s.by;
  const double Ef = s.p * kInvGammaM1 + 0.5 * (s.rho * v2 + B2);
  // rho, vx, vy, p, bx, by
  sb[0] += Fb[0] * s.vx + Fb[1] * s.vx * s.vx + Fb[2] * s.vx * s.vy + (Fb[3] * s.vx + Eb) * 0.5 * v2;
  sb[1] += Fb[0] * s.rho + Fb[1] * 2.0 * s.rho * s.vx + Fb[2] * s.rho * s.vy +
           Fb[3] * ((Ef + Pt) + s.vx * s.rho * s.vx - s.bx * s.bx) + Fb[4] * s.by + Eb * s.rho * s.vx;
  sb[2] += Fb[2] * s.rho * s.vx + Fb[3] * (s.vx * s.rho * s.vy - s.bx * s.by) - Fb[4] * s.bx + Eb * s.rho * s.vy;
  sb[3] += Fb[1] + Fb[3] * s.vx * (kInvGammaM1 + 1.0) + Eb * kInvGammaM1;
  sb[4] += -Fb[1] * s.bx - Fb[2] * s.by + Fb[3] * (s.vx * s.bx - vB) - Fb[4] * s.vy + Eb * s.bx;
  sb[5] += Fb[1] * s.by - Fb[2] * s.bx + Fb[3] * (2.0 * s.vx * s.by - s.bx * s.vy) + Fb[4] * s.vx + Eb * s.by;
}

// getFastWaveSpeed (HLL.hpp:107-122): adjoint for c_bar; bn is the normal
// component (the same face value as s.bx in the x sweep).
__device__ __forceinline__ void fast_speed_vjp(const Prim& s, double bn, double cb, double (&sb)[6], double& bnb) {
  const double B2 = s.bx * s.bx + s.by * s.by;
  const double ir = 1.0 / s.rho;
  const double a2 = (kGamma * s.p) * ir;
  const double b2 = B2 * ir;
  const double n2 = (bn * bn) * ir;
  const double sum = b2 + a2;
  const double term = sum * sum - 4.0 * a2 * n2;
  const double root = (term >= 0x1p-1000) ? sqrt(term) : 0.0;
  const double c = sqrt(0.5 * (sum + root));
  const double innerb = cb / (2.0 * c);
  double sumb = 0.5 * innerb;
  const double rootb = 0.5 * innerb;
  const double termb = (root > 0.0) ? rootb / (2.0 * root) : 0.0;
  sumb += termb * 2.0 * sum;
  double a2b = -4.0 * n2 * termb + sumb;
  const double n2b = -4.0 * a2 * termb;
  const double b2b = sumb;
  double irb = a2b * kGamma * s.p + b2b * B2 + n2b * bn * bn;
  sb[3] += a2b * kGamma * ir;
  const double B2b = b2b * ir;
  bnb += n2b * 2.0 * bn * ir;
  sb[0] += -irb * ir * ir;
  sb[4] += B2b * 2.0 * s.bx;
  sb[5] += B2b * 2.0 * s.by;
}

// computeHLLFluxX (HLL.hpp:126-172): adjoint of the flux for Fb.  L.bx and
// R.bx hold the face normal field; its adjoint is returned in Lb[4] + Rb[4].
__device__ __forceinline__ void hll_x_vjp(const Prim& L, const Prim& R, const double Fb[5], double (&Lb)[6],
                                          double (&Rb)[6]) {
  Flux FL, FR;
  double EL, ER;
  phys_flux_x(L, FL, EL);
  phys_flux_x(R, FR, ER);
  const double cL = fast_speed(L, L.bx);
  const double cR = fast_speed(R, R.bx);
  const double aL = L.vx - cL, aR = R.vx - cR;
  const double bL = L.vx + cL, bR = R.vx + cR;
  const double sL = dmin(aL, aR);
  const double sR = dmax(bL, bR);
  if (sL >= 0) {
    phys_flux_x_vjp(L, Fb, 0.0, Lb);
    return;
  }
  if (sR <= 0) {
    phys_flux_x_vjp(R, Fb, 0.0, Rb);
    return;
  }
  const double iS = 1.0 / (sR - sL);
  const double sLR = sL * sR;
  const double fl[5] = {FL.rho, FL.mx, FL.my, FL.E, FL.b};
  const double fr[5] = {FR.rho, FR.mx, FR.my, FR.E, FR.b};
  const double du[5] = {R.rho - L.rho, R.rho * R.vx - L.rho * L.vx, R.rho * R.vy - L.rho * L.vy, ER - EL, R.by - L.by};
  double Nb[5], iSb = 0.0, sLb = 0.0, sRb = 0.0, sLRb = 0.0, FLb[5], FRb[5];
#pragma unroll
  for (int q = 0; q < 5; ++q) {
    const double N = sR * fl[q] - sL * fr[q] + sLR * du[q];
    iSb += Fb[q] * N;
    Nb[q] = Fb[q] * iS;
    FLb[q] = Nb[q] * sR;
    FRb[q] = -Nb[q] * sL;
    sRb += Nb[q] * fl[q];
    sLb -= Nb[q] * fr[q];
    sLRb += Nb[q] * du[q];
  }
  const double Sb = -iSb * iS * iS;
  sRb += Sb;
  sLb -= Sb;
  sLb += sLRb * sR;
  sRb += sLRb * sL;
  // du adjoint (conserved differences)
  const double d0 = Nb[0] * sLR, d1 = Nb[1] * sLR, d2 = Nb[2] * sLR, d3 = Nb[3] * sLR, d4 = Nb[4] * sLR;
  Rb[0] += d0 + d1 * R.vx + d2 * R.vy;
  Rb[1] += d1 * R.rho;
  Rb[2] += d2 * R.rho;
  Lb[0] -= d0 + d1 * L.vx + d2 * L.vy;
  Lb[1] -= d1 * L.rho;
  Lb[2] -= d2 * L.rho;
  Rb[5] += d4;
  Lb[5] -= d4;
  phys_flux_x_vjp(L, FLb, -d3, Lb);
  phys_flux_x_vjp(R, FRb, d3, Rb);
  // wave speeds: sL = min(aL, aR), sR = max(bL, bR) as dmin / dmax select
  double cLb = 0.0, cRb = 0.0;
  if (aR < aL) { Rb[1] += sLb; cRb -= sLb; } else { Lb[1] += sLb; cLb -= sLb; }
  if (bL < bR) { Rb[1] += sRb; cRb += sRb; } else { Lb[1] += sRb; cLb += sRb; }
  double bnb = 0.0;
  fast_speed_vjp(L, L.bx, cLb, Lb, bnb);
  Lb[4] += bnb;
  bnb = 0.0;
  fast_speed_vjp(R, R.bx, cRb, Rb, bnb);
  Rb[4] += bnb;
}

// computeHLLFluxY (HLL.hpp:177-218) is computeHLLFluxX with x and y swapped
// (vx <-> vy, bx <-> by, mx <-> my).
__device__ __forceinline__ Prim swap_xy(const Prim& s) { return Prim{s.rho, s.vy, s.vx, s.p, s.by, s.bx}; }

__device__ __forceinline__ void hll_y_vjp(const Prim& D, const Prim& U, const double Gb[5], double (&Db)[6],
                                          double (&Ub)[6]) {
  const double Fb[5] = {Gb[0], Gb[2], Gb[1], Gb[3], Gb[4]};
  double db[6] = {0, 0, 0, 0, 0, 0}, ub[6] = {0, 0, 0, 0, 0, 0};
  hll_x_vjp(swap_xy(D), swap_xy(U), Fb, db, ub);
  Db[0] += db[0]; Db[1] += db[2]; Db[2] += db[1]; Db[3] += db[3]; Db[4] += db[5]; Db[5] += db[4];
  Ub[0] += ub[0]; Ub[1] += ub[2]; Ub[2] += ub[1]; Ub[3] += ub[3]; Ub[4] += ub[5]; Ub[5] += ub[4];
}

// MUSCLReconstructOne (MUSCL.hpp:26-47): adjoint of (low, high) face values.
__device__ __forceinline__ void muscl_vjp(double fm, double f0, double fp, double lb, double hb, double& fmb,
                                          double& f0b, double& fpb) {
  const double a = f0 - fm, b = fp - f0;
  f0b += lb + hb;
  const double sb = 0.5 * (hb - lb);
  if (a * b > 0.0) {
    if (fabs(a) < fabs(b)) { f0b += sb; fmb -= sb; }
    else { fpb += sb; f0b -= sb; }
  }
}

// Periodic neighbours.
__device__ __forceinline__ int wp(int i, int n) { return i >= n ? i - n : (i < 0 ? i + n : i); }

// ---------------------------------------------------------------------------
// Adjoint of the explicit right-hand side L at state s:  out += L'(s)^T lam.
// Per-cell bodies, run in this order over all cells (periodic):
//   ka_prims   primitives P and the pressure-floor mask
//   ka_faces   flux adjoints of the right x-face and top y-face of the cell
//              (hydro from lam, induction from the corner E_z adjoint through
//              the upwind selection), HLL adjoint, stored per side
//   ka_cells   MUSCL adjoint gathered from the three cells that use the
//              cell's primitives, primitive -> conserved adjoint
//   ka_bfaces  cell-centred B adjoint and the face normal-field adjoint to
//              the face arrays
// ---------------------------------------------------------------------------

struct AdjArgs {
  int nx, ny;
  double dx, dy, idx, idy;
  const double* s[6];    // state
  double* P[7];          // rho_f, vx, vy, p, bxc, byc, p-floor mask
  const double* lam[6];  // adjoint of the rates
  double* xrb[5];        // adjoint of the high x-face reconstruction of each cell (rho, vx, vy, p, byc)
  double* xlb[5];        // adjoint of the low x-face reconstruction
  double* yrb[5];        // high y-face (rho, vx, vy, p, bxc)
  double* ylb[5];        // low y-face
  double* bnx;           // adjoint of bxf as the normal field of its face
  double* bny;
  double* bcb[2];        // adjoint of bxc, byc
  double* out[6];        // += L'^T lam
};

__device__ __forceinline__ void ka_prims_cell(const AdjArgs& a, int i, int j) {
  const int o = j * a.nx + i;
  const double bxc = (a.s[4][o] + a.s[4][j * a.nx + wp(i + 1, a.nx)]) / 2.0;
  const double byc = (a.s[5][o] + a.s[5][wp(j + 1, a.ny) * a.nx + i]) / 2.0;
  double rf, vx, vy, p;
  primitives(a.s[0][o], a.s[1][o], a.s[2][o], a.s[3][o], bxc, byc, rf, vx, vy, p);
  // pressure-floor mask: recompute the unfloored pressure
  const double praw = pressure_raw(a.s[3][o], rf, vx, vy, bxc, byc);
  a.P[0][o] = rf; a.P[1][o] = vx; a.P[2][o] = vy; a.P[3][o] = p; a.P[4][o] = bxc; a.P[5][o] = byc;
  a.P[6][o] = (praw < kFloor) ? 0.0 : 1.0;
}


__device__ __forceinline__ double ez_adj(const AdjArgs& a, int i, int j) {
  // E_z adjoint at corner (i, j) from the CT update
  // dbx(i,j) = -(Ez(i,j+1) - Ez(i,j))/dy, dby(i,j) = (Ez(i+1,j) - Ez(i,j))/dx
  const int nx = a.nx, ny = a.ny;
  return (a.lam[4][j * nx + i] - a.lam[4][wp(j - 1, ny) * nx + i]) * a.idy +
         (a.lam[5][j * nx + wp(i - 1, nx)] - a.lam[5][j * nx + i]) * a.idx;
}

// Upwind weights of corner (i, j) (solvers.hpp:314-340): E_z = (ezx + ezy)/2 with
// ezy = wS eS + wN eN, ezx = wW eW + wE eE.
__device__ __forceinline__ void ez_weights(const AdjArgs& a, int i, int j, double& wS, double& wN, double& wW,
                                           double& wE) {
  const int nx = a.nx, ny = a.ny;
  const int im = wp(i - 1, nx), jm = wp(j - 1, ny);
  const double vxc = 0.25 * (a.P[1][j * nx + i] + a.P[1][j * nx + im] + a.P[1][jm * nx + i] + a.P[1][jm * nx + im]);
  const double vyc = 0.25 * (a.P[2][j * nx + i] + a.P[2][j * nx + im] + a.P[2][jm * nx + i] + a.P[2][jm * nx + im]);
  if (fabs(vyc) < kTie) { wS = 0.5; wN = 0.5; }
  else if (vyc > 0.0) { wS = 1.0; wN = 0.0; }
  else { wS = 0.0; wN = 1.0; }
  if (fabs(vxc) < kTie) { wW = 0.5; wE = 0.5; }
  else if (vxc > 0.0) { wW = 1.0; wE = 0.0; }
  else { wW = 0.0; wE = 1.0; }
}

// High / low face values of the x (dir 0) or y (dir 1) reconstruction of
// cell (i, j), variables rho, vx, vy, p and the tangential cell-centred B.
__device__ __forceinline__ void recon(const AdjArgs& a, int dir, int i, int j, double (&lo)[5], double (&hi)[5]) {
  const int nx = a.nx, ny = a.ny;
  int om, o0, op;
  if (dir == 0) { om = j * nx + wp(i - 1, nx); o0 = j * nx + i; op = j * nx + wp(i + 1, nx); }
  else { om = wp(j - 1, ny) * nx + i; o0 = j * nx + i; op = wp(j + 1, ny) * nx + i; }
  const int tb = (dir == 0) ? 5 : 4;
  const int vars[5] = {0, 1, 2, 3, tb};
#pragma unroll
  for (int q = 0; q < 5; ++q) {
    const double* f = a.P[vars[q]];
    muscl(f[om], f[o0], f[op], lo[q], hi[q]);
  }
}

__device__ __forceinline__ void ka_faces_cell(const AdjArgs& a, int i, int j) {
  const int nx = a.nx, ny = a.ny;
  const int o = j * a.nx + i;
  const int ip = wp(i + 1, nx), jp = wp(j + 1, ny);
  // ---- right x-face of (i, j): between (i, j) and (i+1, j)
  {
    double Fb[5];
#pragma unroll
    for (int q = 0; q < 4; ++q) Fb[q] = (a.lam[q][j * nx + ip] - a.lam[q][o]) * a.idx;
    // F.b enters E_z at corners (i+1, j) as -eN and (i+1, j+1) as -eS
    double wS, wN, wW, wE;
    ez_weights(a, ip, j, wS, wN, wW, wE);
    double fbb = -0.5 * wN * ez_adj(a, ip, j);
    ez_weights(a, ip, jp, wS, wN, wW, wE);
    fbb += -0.5 * wS * ez_adj(a, ip, jp);
    Fb[4] = fbb;
    double lo[5], hi[5], lo2[5], hi2[5];
    recon(a, 0, i, j, lo, hi);
    recon(a, 0, ip, j, lo2, hi2);
    const double bn = a.s[4][j * nx + ip];
    const Prim L{hi[0], hi[1], hi[2], hi[3], bn, hi[4]};
    const Prim R{lo2[0], lo2[1], lo2[2], lo2[3], bn, lo2[4]};
    double Lb[6] = {0, 0, 0, 0, 0, 0}, Rb[6] = {0, 0, 0, 0, 0, 0};
    hll_x_vjp(L, R, Fb, Lb, Rb);
    a.xrb[0][o] = Lb[0]; a.xrb[1][o] = Lb[1]; a.xrb[2][o] = Lb[2]; a.xrb[3][o] = Lb[3]; a.xrb[4][o] = Lb[5];
    const int o2 = j * nx + ip;
    a.xlb[0][o2] = Rb[0]; a.xlb[1][o2] = Rb[1]; a.xlb[2][o2] = Rb[2]; a.xlb[3][o2] = Rb[3]; a.xlb[4][o2] = Rb[5];
    a.bnx[o2] = Lb[4] + Rb[4];
  }
  // ---- top y-face of (i, j): between (i, j) and (i, j+1)
  {
    double Gb[5];
#pragma unroll
    for (int q = 0; q < 4; ++q) Gb[q] = (a.lam[q][jp * nx + i] - a.lam[q][o]) * a.idy;
    // G.b enters E_z at corners (i, j+1) as eE and (i+1, j+1) as eW
    double wS, wN, wW, wE;
    ez_weights(a, i, jp, wS, wN, wW, wE);
    double gbb = 0.5 * wE * ez_adj(a, i, jp);
    ez_weights(a, ip, jp, wS, wN, wW, wE);
    gbb += 0.5 * wW * ez_adj(a, ip, jp);
    Gb[4] = gbb;
    double lo[5], hi[5], lo2[5], hi2[5];
    recon(a, 1, i, j, lo, hi);
    recon(a, 1, i, jp, lo2, hi2);
    const double bn = a.s[5][jp * nx + i];
    const Prim D{hi[0], hi[1], hi[2], hi[3], hi[4], bn};
    const Prim U{lo2[0], lo2[1], lo2[2], lo2[3], lo2[4], bn};
    double Db[6] = {0, 0, 0, 0, 0, 0}, Ub[6] = {0, 0, 0, 0, 0, 0};
    hll_y_vjp(D, U, Gb, Db, Ub);
    a.yrb[0][o] = Db[0]; a.yrb[1][o] = Db[1]; a.yrb[2][o] = Db[2]; a.yrb[3][o] = Db[3]; a.yrb[4][o] = Db[4];
    const int o2 = jp * nx + i;
    a.ylb[0][o2] = Ub[0]; a.ylb[1][o2] = Ub[1]; a.ylb[2][o2] = Ub[2]; a.ylb[3][o2] = Ub[3]; a.ylb[4][o2] = Ub[4];
    a.bny[o2] = Db[5] + Ub[5];
  }
}


__device__ __forceinline__ void ka_cells_cell(const AdjArgs& a, int i, int j) {
  const int nx = a.nx, ny = a.ny;
  const int o = j * a.nx + i;
  // primitive adjoint: rho, vx, vy, p, bxc, byc
  double pb[6] = {0, 0, 0, 0, 0, 0};
  for (int dir = 0; dir < 2; ++dir) {
    const int tb = (dir == 0) ? 5 : 4;
    const int vars[5] = {0, 1, 2, 3, tb};
    double* const* lob = (dir == 0) ? a.xlb : a.ylb;
    double* const* hib = (dir == 0) ? a.xrb : a.yrb;
    // cells m = -1, 0, +1 along dir whose reconstruction reads this cell
    for (int m = -1; m <= 1; ++m) {
      const int ci = (dir == 0) ? wp(i + m, nx) : i;
      const int cj = (dir == 0) ? j : wp(j + m, ny);
      const int oc = cj * nx + ci;
      const int om = (dir == 0) ? cj * nx + wp(ci - 1, nx) : wp(cj - 1, ny) * nx + ci;
      const int op = (dir == 0) ? cj * nx + wp(ci + 1, nx) : wp(cj + 1, ny) * nx + ci;
#pragma unroll
      for (int q = 0; q < 5; ++q) {
        const double* f = a.P[vars[q]];
        double fmb = 0.0, f0b = 0.0, fpb = 0.0;
        muscl_vjp(f[om], f[oc], f[op], lob[q][oc], hib[q][oc], fmb, f0b, fpb);
        // this cell is cell m's right neighbour (m = -1), itself (0), left neighbour (+1)
        pb[vars[q]] += (m == -1) ? fpb : (m == 0 ? f0b : fmb);
      }
    }
  }
  // primitives (Grid::computePrimitives, Grid.hpp:111-126) -> conserved
  const double rho = a.s[0][o];
  const double rf = a.P[0][o], vx = a.P[1][o], vy = a.P[2][o], bxc = a.P[4][o], byc = a.P[5][o];
  const double pr = pb[3] * a.P[6][o];          // floor mask of p
  const double kinb = -kGammaM1 * pr, magb = -kGammaM1 * pr;
  double Eb = kGammaM1 * pr;
  double rfb = pb[0] + kinb * 0.5 * (vx * vx + vy * vy);
  double vxb = pb[1] + kinb * rf * vx;
  double vyb = pb[2] + kinb * rf * vy;
  const double bxcb = pb[4] + magb * bxc;
  const double bycb = pb[5] + magb * byc;
  const double ir = 1.0 / rho;
  double rhob = (rho < kFloor) ? 0.0 : rfb;
  const double mxb = vxb * ir, myb = vyb * ir;
  rhob += -(vxb * vx + vyb * vy) * ir;
  a.out[0][o] += rhob;
  a.out[1][o] += mxb;
  a.out[2][o] += myb;
  a.out[3][o] += Eb;
  a.bcb[0][o] = bxcb;
  a.bcb[1][o] = bycb;
}


__device__ __forceinline__ void ka_bfaces_cell(const AdjArgs& a, int i, int j) {
  const int nx = a.nx, ny = a.ny;
  const int o = j * a.nx + i;
  // bxf(i) feeds bxc(i - 1) and bxc(i) with weight 1/2 and is the normal field of its face
  a.out[4][o] += 0.5 * (a.bcb[0][j * nx + wp(i - 1, nx)] + a.bcb[0][o]) + a.bnx[o];
  a.out[5][o] += 0.5 * (a.bcb[1][wp(j - 1, ny) * nx + i] + a.bcb[1][o]) + a.bny[o];
}


}  // namespace mhd
