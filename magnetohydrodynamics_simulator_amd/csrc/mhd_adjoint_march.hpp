// mhd_adjoint_march.hpp -- the adjoint of the explicit right-hand side,
// out += L'(s)^T lam, in one pass: primitives, face adjoints (HLL x / y with
// the upwind E_z selection), MUSCL adjoint, primitive -> conserved adjoint and
// the face-field terms, fused the way the forward stage kernel fuses the
// forward pass.
//
// One wavefront owns a strip of 64 columns, lane l <-> column x0 - 4 + l, and
// marches up a segment of rows; lanes 4..60 produce output (57 columns per
// strip), the others are the halo the stencil needs.  Row k of the march
// computes, with everything older kept in registers:
//   P(k)         primitives and pressure-floor mask of row k
//   C(k)         corner velocities -> upwind weights, E_z adjoint of the
//                bottom-left corners of row k (lam of rows k-1, k)
//   X(k-1)       x-face adjoint right of the cells of row k-1 (HLL x vjp):
//                hi-x adjoint of the cell, lo-x adjoint of its right neighbour
//   MX(k-1)      MUSCL-x adjoint of row k-1 -> primitive adjoint pb(k-1)
//   Y(k-2)       y-face adjoint on top of row k-2 (HLL y vjp)
//   MY(k-2)      MUSCL-y adjoint of row k-2 -> pb(k-3), pb(k-2), pb(k-1)
//   out(k-3)     primitive -> conserved adjoint and the face-field terms
// so output row k-3 is final once row k has been loaded; a segment of H rows
// takes H + 7 row steps.  Every per-face and per-cell formula and every
// summation order is that of the unfused bodies in mhd_adjoint_math.hpp
// (ka_prims / ka_faces / ka_cells / ka_bfaces), which stay as the reference
// the tests compare against.
#pragma once

#include "mhd_adjoint_math.hpp"
#include "mhd_device.hpp"

namespace mhd {

#ifndef MHD_MARCH_WAVES
#define MHD_MARCH_WAVES 2      // waves per SIMD the march kernels are compiled for (tuning parameter)
#endif

constexpr int kAdjHalo = 4;    // lanes left of the first output column
constexpr int kAdjOut = 57;    // output columns per 64-lane strip

struct MarchArgs {
  const double* s[6];    // state
  const double* lam[6];  // adjoint of the rates
  double* out[6];        // += L'^T lam
  int nx, ny, rows;      // rows per segment
  double idx, idy;
};

__device__ __forceinline__ double lane_up(double v) { return __shfl_up(v, 1, 64); }     // value of lane l-1
__device__ __forceinline__ double lane_dn(double v) { return __shfl_down(v, 1, 64); }   // value of lane l+1

// Upwind weights of a corner (ez_weights of mhd_adjoint_math.hpp): wN = 1 - wS
// and wE = 1 - wW exactly (the weights are 0, 1/2 or 1), so two are kept.
struct UpwindW {
  double wS, wW;
  __device__ double wN() const { return 1.0 - wS; }
  __device__ double wE() const { return 1.0 - wW; }
};

__device__ __forceinline__ UpwindW upwind_weights(double vxc, double vyc) {
  UpwindW w;
  w.wS = (fabs(vyc) < kTie) ? 0.5 : ((vyc > 0.0) ? 1.0 : 0.0);
  w.wW = (fabs(vxc) < kTie) ? 0.5 : ((vxc > 0.0) ? 1.0 : 0.0);
  return w;
}

// Primitive row: rho (floored), vx, vy, p, bxc, byc.
struct PRow {
  double v[6];
};

__device__ __forceinline__ void adj_march(const MarchArgs& a, int strip, int seg) {
  const int lane = threadIdx.x;
  const int nx = a.nx, ny = a.ny;
  const int x0 = strip * kAdjOut;
  const int y0 = seg * a.rows;
  const int y1 = min(y0 + a.rows, ny);
  const int col = x0 - kAdjHalo + lane;
  const int c0 = pmod(col, nx);
  const int c1 = pmod(col + 1, nx);
  const bool out_lane = (lane >= kAdjHalo) && (lane < kAdjHalo + kAdjOut) && (col < nx);
  const double idx = a.idx, idy = a.idy;

  // carried state (rows relative to the current row k)
  PRow p1 = {}, p2 = {}, p3 = {};              // P(k-1), P(k-2), P(k-3)
  double lam4_1 = 0.0;                         // lam_4 of row k-1
  UpwindW w1 = {0, 0};                         // corner (c, k-1)
  double ez1 = 0.0;
  double bxr1 = 0.0;                           // bxf(c+1, k-1)
  double byf1 = 0.0;                           // byf(c, k-1)
  double byf0 = a.s[5][pmod(y0 - 4, ny) * nx + c0];    // byf(c, k)
  double yl2[5] = {0, 0, 0, 0, 0};             // lo-y adjoint of cell (c, k-2) (y-face under it)
  double bny2 = 0.0, bny3 = 0.0;               // face-normal adjoint of byf(c, k-2), byf(c, k-3)
  double bnx2 = 0.0, bnx3 = 0.0;               // face-normal adjoint of bxf(c+1, k-2), bxf(c+1, k-3)
  double acc2[6] = {0, 0, 0, 0, 0, 0}, acc3[6] = {0, 0, 0, 0, 0, 0};   // pb of rows k-2, k-3 so far
  double bycb4 = 0.0;                          // byc adjoint of row k-4

  for (int k = y0 - 4; k <= y1 + 2; ++k) {
    // rows wrap periodically (the march reaches 7 rows below a segment,
    // more than one period when ny < 7)
    const int rk = pmod(k, ny) * nx;
    const int rk1 = pmod(k + 1, ny) * nx;
    const int rm1 = pmod(k - 1, ny) * nx;
    const int rm2 = pmod(k - 2, ny) * nx;
    // ---- P(k) ----
    PRow p0;
    const double bxr0 = a.s[4][rk + c1];
    const double byf_n = a.s[5][rk1 + c0];
    {
      const double bxc = (a.s[4][rk + c0] + bxr0) / 2.0;
      const double byc = (byf0 + byf_n) / 2.0;
      double rf, vx, vy, p;
      primitives(a.s[0][rk + c0], a.s[1][rk + c0], a.s[2][rk + c0], a.s[3][rk + c0], bxc, byc, rf, vx, vy, p);
      p0.v[0] = rf; p0.v[1] = vx; p0.v[2] = vy; p0.v[3] = p; p0.v[4] = bxc; p0.v[5] = byc;
    }
    // lam_0..3 of rows k-1 and k-2 are read again (cache hits) rather than
    // carried: registers are the limit of this kernel
    double lam1[4], lam2[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      lam1[q] = a.lam[q][rm1 + c0];
      lam2[q] = a.lam[q][rm2 + c0];
    }
    const double lam4_0 = a.lam[4][rk + c0];
    const double lam5_0 = a.lam[5][rk + c0];

    // ---- C(k): corner (c, k) ----
    UpwindW w0;
    double ez0;
    {
      const double vxc = 0.25 * (p0.v[1] + lane_up(p0.v[1]) + p1.v[1] + lane_up(p1.v[1]));
      const double vyc = 0.25 * (p0.v[2] + lane_up(p0.v[2]) + p1.v[2] + lane_up(p1.v[2]));
      w0 = upwind_weights(vxc, vyc);
      ez0 = (lam4_0 - lam4_1) * idy + (lane_up(lam5_0) - lam5_0) * idx;
    }

    // ---- X(k-1): x-face between (c, k-1) and (c+1, k-1) ----
    double xh[5], xl_next[5], bnx1;
    {
      double Fb[5];
#pragma unroll
      for (int q = 0; q < 4; ++q) Fb[q] = (lane_dn(lam1[q]) - lam1[q]) * idx;
      double fbb = -0.5 * lane_dn(w1.wN()) * lane_dn(ez1);
      fbb += -0.5 * lane_dn(w0.wS) * lane_dn(ez0);
      Fb[4] = fbb;
      const int xv[5] = {0, 1, 2, 3, 5};
      double lo[5], hi[5];
#pragma unroll
      for (int q = 0; q < 5; ++q) muscl(lane_up(p1.v[xv[q]]), p1.v[xv[q]], lane_dn(p1.v[xv[q]]), lo[q], hi[q]);
      const double bn = bxr1;
      const Prim L{hi[0], hi[1], hi[2], hi[3], bn, hi[4]};
      const Prim R{lane_dn(lo[0]), lane_dn(lo[1]), lane_dn(lo[2]), lane_dn(lo[3]), bn, lane_dn(lo[4])};
      double Lb[6] = {0, 0, 0, 0, 0, 0}, Rb[6] = {0, 0, 0, 0, 0, 0};
      hll_x_vjp(L, R, Fb, Lb, Rb);
      xh[0] = Lb[0]; xh[1] = Lb[1]; xh[2] = Lb[2]; xh[3] = Lb[3]; xh[4] = Lb[5];
      xl_next[0] = Rb[0]; xl_next[1] = Rb[1]; xl_next[2] = Rb[2]; xl_next[3] = Rb[3]; xl_next[4] = Rb[5];
      bnx1 = Lb[4] + Rb[4];
    }

    // ---- MX(k-1): pb(k-1) from the MUSCL-x adjoints of cells c-1, c, c+1 ----
    double acc1[6] = {0, 0, 0, 0, 0, 0};
    {
      const int xv[5] = {0, 1, 2, 3, 5};
#pragma unroll
      for (int q = 0; q < 5; ++q) {
        const double f = p1.v[xv[q]];
        double fmb = 0.0, f0b = 0.0, fpb = 0.0;
        muscl_vjp(lane_up(f), f, lane_dn(f), lane_up(xl_next[q]), xh[q], fmb, f0b, fpb);
        acc1[xv[q]] += lane_up(fpb);
        acc1[xv[q]] += f0b;
        acc1[xv[q]] += lane_dn(fmb);
      }
    }

    // ---- Y(k-2): y-face between (c, k-2) and (c, k-1) ----
    double yh[5], yl1[5], bny1;
    {
      double Gb[5];
#pragma unroll
      for (int q = 0; q < 4; ++q) Gb[q] = (lam1[q] - lam2[q]) * idy;
      double gbb = 0.5 * w1.wE() * ez1;
      gbb += 0.5 * lane_dn(w1.wW) * lane_dn(ez1);
      Gb[4] = gbb;
      const int yv[5] = {0, 1, 2, 3, 4};
      double lo[5], hi[5], lo2[5], hi2[5];
#pragma unroll
      for (int q = 0; q < 5; ++q) {
        muscl(p3.v[yv[q]], p2.v[yv[q]], p1.v[yv[q]], lo[q], hi[q]);
        muscl(p2.v[yv[q]], p1.v[yv[q]], p0.v[yv[q]], lo2[q], hi2[q]);
      }
      const double bn = byf1;
      const Prim D{hi[0], hi[1], hi[2], hi[3], hi[4], bn};
      const Prim U{lo2[0], lo2[1], lo2[2], lo2[3], lo2[4], bn};
      double Db[6] = {0, 0, 0, 0, 0, 0}, Ub[6] = {0, 0, 0, 0, 0, 0};
      hll_y_vjp(D, U, Gb, Db, Ub);
      yh[0] = Db[0]; yh[1] = Db[1]; yh[2] = Db[2]; yh[3] = Db[3]; yh[4] = Db[4];
      yl1[0] = Ub[0]; yl1[1] = Ub[1]; yl1[2] = Ub[2]; yl1[3] = Ub[3]; yl1[4] = Ub[4];
      bny1 = Db[5] + Ub[5];
    }

    // ---- MY(k-2): MUSCL-y adjoint of cell (c, k-2) -> rows k-3, k-2, k-1 ----
    {
      const int yv[5] = {0, 1, 2, 3, 4};
#pragma unroll
      for (int q = 0; q < 5; ++q) {
        double fmb = 0.0, f0b = 0.0, fpb = 0.0;
        muscl_vjp(p3.v[yv[q]], p2.v[yv[q]], p1.v[yv[q]], yl2[q], yh[q], fmb, f0b, fpb);
        acc1[yv[q]] += fpb;
        acc2[yv[q]] += f0b;
        acc3[yv[q]] += fmb;
      }
    }

    // ---- out(k-3): primitive -> conserved adjoint, face-field terms ----
    {
      const int j = k - 3;
      const int rj = pmod(j, ny) * nx;
      const double* pb = acc3;
      const double rho = a.s[0][rj + c0];
      const double E = a.s[3][rj + c0];
      const double rf = p3.v[0], vx = p3.v[1], vy = p3.v[2], bxc = p3.v[4], byc = p3.v[5];
      // pressure-floor mask (ka_prims_cell)
      const double mask = (pressure_raw(E, rf, vx, vy, bxc, byc) < kFloor) ? 0.0 : 1.0;
      const double pr = pb[3] * mask;
      const double kinb = -kGammaM1 * pr, magb = -kGammaM1 * pr;
      const double Eb = kGammaM1 * pr;
      const double rfb = pb[0] + kinb * 0.5 * (vx * vx + vy * vy);
      const double vxb = pb[1] + kinb * rf * vx;
      const double vyb = pb[2] + kinb * rf * vy;
      const double bxcb = pb[4] + magb * bxc;
      const double bycb = pb[5] + magb * byc;
      const double ir = 1.0 / rho;
      double rhob = (rho < kFloor) ? 0.0 : rfb;
      const double mxb = vxb * ir, myb = vyb * ir;
      rhob += -(vxb * vx + vyb * vy) * ir;
      const double bx_left = lane_up(bxcb);
      const double bnx_here = lane_up(bnx3);
      if (j >= y0 && j < y1 && out_lane) {
        const int o = rj + c0;
        double ov[6];
#pragma unroll
        for (int q = 0; q < 6; ++q) ov[q] = a.out[q][o];
        a.out[0][o] = ov[0] + rhob;
        a.out[1][o] = ov[1] + mxb;
        a.out[2][o] = ov[2] + myb;
        a.out[3][o] = ov[3] + Eb;
        a.out[4][o] = ov[4] + (0.5 * (bx_left + bxcb) + bnx_here);
        a.out[5][o] = ov[5] + (0.5 * (bycb4 + bycb) + bny3);
      }
      bycb4 = bycb;
    }

    // ---- rotate the carried state ----
    p3 = p2; p2 = p1; p1 = p0;
    lam4_1 = lam4_0;
    w1 = w0;
    ez1 = ez0;
    bxr1 = bxr0;
    byf1 = byf0;
    byf0 = byf_n;
#pragma unroll
    for (int q = 0; q < 5; ++q) yl2[q] = yl1[q];
    bny3 = bny2; bny2 = bny1;
    bnx3 = bnx2; bnx2 = bnx1;
#pragma unroll
    for (int q = 0; q < 6; ++q) { acc3[q] = acc2[q]; acc2[q] = acc1[q]; }
  }
}

}  // namespace mhd
