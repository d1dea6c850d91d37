// mhd_adjoint_lin.hpp -- per-cell bodies of the linear pieces of the IMEX
// step's adjoint: the resistive stencils and their transposes, the inner
// products of the eta derivative, the stage entry/exit combinations and the
// LDS-tiled Jacobi sweeps of the adjoint solves.  The single-run kernels
// (mhd_adjoint.hip) run them over one grid, the batched kernels
// (mhd_batch_adjoint.hip) over the runs of a batch.
//
// Stencils (periodic, corner x at (i, j) = lower-left corner of cell (i, j)):
//   J(B)(i,j)  = (by(i,j) - by(i-1,j))/dx - (bx(i,j) - bx(i,j-1))/dy   (corner)
//   R(x)       = (-(x(i,j+1) - x(i,j))/dy, (x(i+1,j) - x(i,j))/dx)    (faces)
//   Lap(x)     = (x(i-1,j) - 2x + x(i+1,j))/dx^2 + (x(i,j-1) - 2x + x(i,j+1))/dy^2
// J is the corner current of solvers.hpp:450-463, R the face rate -curl_z of
// solvers.hpp:472-488.  A CG solution x whose solve did no product (spmv = 0,
// x = 0 in Eigen) is read as zero: bodies that take a CgState pointer check it.
#pragma once

#include <algorithm>
#include <cmath>

#include "mhd_adjoint_math.hpp"
#include "mhd_device.hpp"
#include "mhd_kernels.hpp"

namespace mhd {

struct LinArgs {
  int nx, ny;
  double idx, idy;
};

__device__ __forceinline__ bool x_is_zero(const CgState* st) { return st != nullptr && st->spmv == 0; }

__device__ __forceinline__ double corner_J(const LinArgs& g, const double* bx, const double* by, int i, int j) {
  const int o = j * g.nx + i;
  return (by[o] - by[j * g.nx + wp(i - 1, g.nx)]) * g.idx - (bx[o] - bx[wp(j - 1, g.ny) * g.nx + i]) * g.idy;
}

// (fx, fy) += c J^T(q)
__device__ __forceinline__ void Jt_cell(const LinArgs& g, const double* q, double c, double* fx, double* fy, int i,
                                        int j) {
  const int o = j * g.nx + i;
  fx[o] += c * (q[wp(j + 1, g.ny) * g.nx + i] - q[o]) * g.idy;
  fy[o] += c * (q[o] - q[j * g.nx + wp(i + 1, g.nx)]) * g.idx;
}

// t = c J^T(q) and out = base + t
__device__ __forceinline__ void Jt2_cell(const LinArgs& g, const double* q, double c, const double* bx,
                                         const double* by, double* tx, double* ty, double* ox, double* oy, int i,
                                         int j) {
  const int o = j * g.nx + i;
  const double ax = c * (q[wp(j + 1, g.ny) * g.nx + i] - q[o]) * g.idy;
  const double ay = c * (q[o] - q[j * g.nx + wp(i + 1, g.nx)]) * g.idx;
  tx[o] = ax;
  ty[o] = ay;
  ox[o] = bx[o] + ax;
  oy[o] = by[o] + ay;
}

// (fx, fy) = (bx, by) + c R(x); zx: x reads as zero
__device__ __forceinline__ void R_cell(const LinArgs& g, const double* bx, const double* by, const double* x,
                                       bool zx, double c, double* fx, double* fy, int i, int j) {
  const int o = j * g.nx + i;
  const double xc = zx ? 0.0 : x[o];
  const double xn = zx ? 0.0 : x[wp(j + 1, g.ny) * g.nx + i];
  const double xe = zx ? 0.0 : x[j * g.nx + wp(i + 1, g.nx)];
  fx[o] = bx[o] + c * (-(xn - xc) * g.idy);
  fy[o] = by[o] + c * ((xe - xc) * g.idx);
}

// q = c R^T(fx, fy)
__device__ __forceinline__ void Rt_cell(const LinArgs& g, const double* fx, const double* fy, double c, double* q,
                                        int i, int j) {
  const int o = j * g.nx + i;
  q[o] = c * ((fx[o] - fx[wp(j - 1, g.ny) * g.nx + i]) * g.idy + (fy[j * g.nx + wp(i - 1, g.nx)] - fy[o]) * g.idx);
}

// u(i,j) Lap(v)(i,j)
__device__ __forceinline__ double lap_dot_cell(const LinArgs& g, const double* u, const double* v, int i, int j) {
  const int o = j * g.nx + i;
  const double lv = (v[j * g.nx + wp(i - 1, g.nx)] - 2.0 * v[o] + v[j * g.nx + wp(i + 1, g.nx)]) * (g.idx * g.idx) +
                    (v[wp(j - 1, g.ny) * g.nx + i] - 2.0 * v[o] + v[wp(j + 1, g.ny) * g.nx + i]) * (g.idy * g.idy);
  return u[o] * lv;
}

// block reduction of acc; thread 0 writes the block's sum to *out
__device__ __forceinline__ void block_partial(double acc, double* out) {
  __shared__ double lds[8];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) lds[w] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = lds[0];
    for (int ww = 1; ww < (int)(blockDim.x >> 6); ++ww) t += lds[ww];
    *out = t;
  }
}

// the 256-thread block sum of v in a fixed order; every thread gets it
__device__ __forceinline__ double block_total(double v, double* lds) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) lds[w] = v;
  __syncthreads();
  double t = lds[0];
  for (int ww = 1; ww < (int)(blockDim.x >> 6); ++ww) t += lds[ww];
  __syncthreads();
  return t;
}

struct Fields6 {
  double* f[6];
};

// Entry of stage 2's adjoint: Wbar = U3*bar with U3*.hbar = lam.h, the density
// floor of U3* as a mask (U3*.rho = the next state's density), fbar = ca Wbar,
// Sbar = 0.
__device__ __forceinline__ void stage2_in_cell(int o, const Fields6& lam, const Fields6& wb, const Fields6& fb,
                                               const Fields6& sb, const double* rho_next, double ca) {
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    double w = (q < 4) ? lam.f[q][o] : wb.f[q][o];
    if (q == 0 && !(rho_next[o] > kFloor)) w = 0.0;
    if (q < 4) wb.f[q][o] = w;
    fb.f[q][o] = ca * w;
    sb.f[q][o] = 0.0;
  }
}

// Exit of stage 1's adjoint: density floor of S* as a mask (S.rho = S*.rho),
// fbar = cs S*bar + cw Wbar, lam = S*bar + Wbar.
__device__ __forceinline__ void stage1_out_cell(int o, const Fields6& sb, const Fields6& wb, const Fields6& fb,
                                                const Fields6& lam, const double* rho_s, double cs, double cw) {
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    double sv = sb.f[q][o];
    if (q == 0 && !(rho_s[o] > kFloor)) sv = 0.0;
    const double wv = wb.f[q][o];
    fb.f[q][o] = cs * sv + cw * wv;
    lam.f[q][o] = sv + wv;
  }
}

// Tile (tx, ty) of S Jacobi sweeps x <- (b + cx (x_w + x_e) + cy (x_s + x_n)) / d
// of A x = b, A = d I - cx (E + W) - cy (N + S) (periodic), on T x T tiles
// staged in LDS with a halo of S corners: each sweep shrinks the region that
// is still exact by one, so the tile's T x T interior is exact after S
// sweeps.  xin == nullptr starts from x = 0.  256 threads; x ping-pongs in
// LDS, b is held in registers (a third less LDS, more blocks per CU).
template <int T, int S>
__device__ __forceinline__ void jacobi_tile(const LinArgs& g, double cx, double cy, double invd, const double* b,
                                            const double* xin, double* xout, int tx, int ty) {
  constexpr int W = T + 2 * S;
  constexpr int kPts = (W * W + 255) / 256;   // points per thread (blockDim.x == 256)
  __shared__ double xs[2][W * W];
  double bv[kPts];                            // b of the thread's points stays in registers
  const int x0 = tx * T - S, y0 = ty * T - S;
#pragma unroll
  for (int m = 0; m < kPts; ++m) {
    const int k = threadIdx.x + 256 * m;
    bv[m] = 0.0;
    if (k < W * W) {
      const int i = k % W, j = k / W;
      const int gi = pmod(x0 + i, g.nx), gj = pmod(y0 + j, g.ny);
      bv[m] = b[gj * g.nx + gi];
      xs[0][k] = xin ? xin[gj * g.nx + gi] : 0.0;
    }
  }
  __syncthreads();
  for (int sw = 0; sw < S; ++sw) {
    const double* cur = xs[sw & 1];
    double* nxt = xs[(sw & 1) ^ 1];
#pragma unroll
    for (int m = 0; m < kPts; ++m) {
      const int k = threadIdx.x + 256 * m;
      const int i = k % W, j = k / W;
      if (k < W * W && i > sw && i < W - 1 - sw && j > sw && j < W - 1 - sw)
        nxt[k] = (bv[m] + cx * (cur[k - 1] + cur[k + 1]) + cy * (cur[k - W] + cur[k + W])) * invd;
    }
    __syncthreads();
  }
  const double* res = xs[S & 1];
  for (int k = threadIdx.x; k < T * T; k += blockDim.x) {
    const int i = k % T, j = k / T;
    const int gi = tx * T + i, gj = ty * T + j;
    if (gi < g.nx && gj < g.ny) xout[gj * g.nx + gi] = res[(j + S) * W + (i + S)];
  }
}

constexpr int kJacTile = 32;    // Jacobi tile side
constexpr int kJacSweeps = 8;   // sweeps per launch (the halo width)

// Jacobi sweeps from x = 0 for A = d I - cx (E + W) - cy (N + S): the max-norm
// error after k sweeps is at most rho^k |x|_inf with the contraction factor
// rho = 2 (cx + cy) / d, so ceil(53 ln 2 / ln(1/rho)) sweeps bring it below
// 2^-53 |x|_inf (half an ulp of the largest entry).  Used where rho <= 1/4
// (the operators of small alpha / h^2); 0 (solve by CG) otherwise.
inline int jacobi_sweep_count(double cx, double cy) {
  const double d = 1 + 2 * (cx + cy);
  const double rho = 2 * (cx + cy) / d;
  if (!(rho <= 0.25)) return 0;
  if (rho < 1e-300) return 1;
  return std::max(1, (int)std::ceil(-53.0 * std::log(2.0) / std::log(rho)));
}

}  // namespace mhd
