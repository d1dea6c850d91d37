// mhd_kernels.hip -- gfx950 kernels of the 2D resistive-MHD ARS(2,2,2) step.
//
// Device layout: every field is a compact, periodic nx * ny array of doubles,
// row-major (index j * nx + i), structure of arrays.  bx[j*nx+i] is B_x on the
// left face x_{i-1/2} of cell (i, j), by[j*nx+i] is B_y on the bottom face
// y_{j-1/2}, corner quantities (E_z, the CG unknowns) sit at the bottom-left
// corner -- the staggering of the reference's Grid (Grid.hpp:19-26,
// solvers.hpp:312), without its ghost layers: periodicity is applied in the
// index arithmetic instead of by ghost-fill passes.
#include <hip/hip_runtime.h>

#include "mhd_device.hpp"
#include "mhd_kernels.hpp"

namespace mhd {

// ---------------------------------------------------------------------------
// Fused explicit stage: primitives -> MUSCL -> HLL x/y -> upwind CT E_z ->
// flux divergence and CT face update -> stage combination, in one pass.
//
// One wavefront owns a strip of 64 columns, lane l <-> column x0 - 2 + l, and
// marches up a segment of rows.  Lanes 2..61 produce output (60 columns per
// strip); lanes 0,1,62,63 are the halo the stencil needs.  Row k of the march
// computes, with everything older kept in registers:
//   P(k)       primitives of row k                     (loads of row k)
//   YR(k-1)    y-reconstruction of row k-1             (P(k-2), P(k-1), P(k))
//   G(k-2)     y-face flux on top of row k-2           (YR(k-2).R, YR(k-1).L)
//   XR(k-1)    x-reconstruction of row k-1             (cross-lane neighbours)
//   F(k-1)     x-face flux right of row k-1 cells      (XR of this and next lane)
//   Ez(k-1)    corner E_z, bottom-left of row k-1      (F(k-1), F(k-2), G(k-2))
//   dU(k-2)    flux divergence + CT of row k-2         (F, G, Ez)
// so output row k-2 is final once row k has been loaded; a segment of H rows
// takes H + 4 row steps.  Neighbouring columns move between lanes with
// cross-lane shuffles; no LDS and no barriers are used.
// ---------------------------------------------------------------------------

struct Carry {
  // primitives of rows k-2 (a) and k-1 (b): rho (floored), vx, vy, p, bxc, byc
  double a_rho, a_vx, a_vy, a_p, a_bxc;
  double b_rho, b_vx, b_vy, b_p, b_bxc, b_byc;
  double yr_rho, yr_vx, yr_vy, yr_p, yr_bx;   // YR(k-2).R
  double f_rho, f_mx, f_my, f_E;              // F(c, k-2)
  double fby_left;                            // F.by(c-1, k-2)
  double g_rho, g_mx, g_my, g_E;              // G(c, k-3)
  double ez;                                  // Ez(c, k-2)
  double vx_left_a, vy_left_a;                // vx, vy of (c-1, k-2)
  double bxr_b;                               // bxf(c+1, k-1)
  double byf_b;                               // byf(c, k-1)
  double byf_c;                               // byf(c, k)
  // resistive explicit term of stage 2: J at (c, k-2) and face B history
  double j_a;
  double bxl_b, bxl_a;                        // bxf(c, k-1), bxf(c, k-2)
};

__device__ __forceinline__ double sh_up(double v) { return __shfl_up(v, 1, 64); }     // value of lane l-1
__device__ __forceinline__ double sh_dn(double v) { return __shfl_down(v, 1, 64); }   // value of lane l+1

template <int MODE>
__global__ __launch_bounds__(64) void k_stage(StageArgs a) {
  const int lane = threadIdx.x;
  const int nx = a.nx, ny = a.ny;
  const int x0 = blockIdx.x * kStripOut;
  const int y0 = blockIdx.y * a.rows;
  const int y1 = min(y0 + a.rows, ny);
  const int col = x0 - 2 + lane;
  const int c0 = pmod(col, nx);
  const int c1 = pmod(col + 1, nx);
  const bool out_lane = (lane >= 2) && (lane < 2 + kStripOut) && (col < nx);
  const double idx_ = a.inv_dx, idy_ = a.inv_dy;

  Carry s = {};
  s.byf_c = a.s[5][wrap1(y0 - 2, ny) * nx + c0];
  s.bxl_b = 0.0;

  for (int k = y0 - 2; k <= y1 + 1; ++k) {
    const int rk = wrap1(k, ny);
    const int rk1 = wrap1(k + 1, ny);
    const int base = rk * nx;
    // ---- loads of row k ----
    const double rho = a.s[0][base + c0];
    const double mx = a.s[1][base + c0];
    const double my = a.s[2][base + c0];
    const double E = a.s[3][base + c0];
    const double bxl = a.s[4][base + c0];
    const double bxr = a.s[4][base + c1];
    const double byn = a.s[5][rk1 * nx + c0];
    // ---- P(k) ----
    const double bxc = (bxl + bxr) / 2.0;
    const double byc = (s.byf_c + byn) / 2.0;
    double c_rho, c_vx, c_vy, c_p;
    primitives(rho, mx, my, E, bxc, byc, c_rho, c_vx, c_vy, c_p);

    // ---- YR(k-1) and G(k-2) ----
    double yl_rho, yr_rho, yl_vx, yr_vx, yl_vy, yr_vy, yl_p, yr_p, yl_bx, yr_bx;
    muscl(s.a_rho, s.b_rho, c_rho, yl_rho, yr_rho);
    muscl(s.a_vx, s.b_vx, c_vx, yl_vx, yr_vx);
    muscl(s.a_vy, s.b_vy, c_vy, yl_vy, yr_vy);
    muscl(s.a_p, s.b_p, c_p, yl_p, yr_p);
    muscl(s.a_bxc, s.b_bxc, bxc, yl_bx, yr_bx);
    Flux G;
    {
      Prim D{s.yr_rho, s.yr_vx, s.yr_vy, s.yr_p, s.yr_bx, s.byf_b};
      Prim U{yl_rho, yl_vx, yl_vy, yl_p, yl_bx, s.byf_b};
      G = hll_y(D, U);
    }

    // ---- XR(k-1) and F(k-1) ----
    const double l_rho = sh_up(s.b_rho), r_rho = sh_dn(s.b_rho);
    const double l_vx = sh_up(s.b_vx), r_vx = sh_dn(s.b_vx);
    const double l_vy = sh_up(s.b_vy), r_vy = sh_dn(s.b_vy);
    const double l_p = sh_up(s.b_p), r_p = sh_dn(s.b_p);
    const double l_by = sh_up(s.b_byc), r_by = sh_dn(s.b_byc);
    double xl_rho, xr_rho, xl_vx, xr_vx, xl_vy, xr_vy, xl_p, xr_p, xl_by, xr_by;
    muscl(l_rho, s.b_rho, r_rho, xl_rho, xr_rho);
    muscl(l_vx, s.b_vx, r_vx, xl_vx, xr_vx);
    muscl(l_vy, s.b_vy, r_vy, xl_vy, xr_vy);
    muscl(l_p, s.b_p, r_p, xl_p, xr_p);
    muscl(l_by, s.b_byc, r_by, xl_by, xr_by);
    Flux F;
    {
      Prim L{xr_rho, xr_vx, xr_vy, xr_p, s.bxr_b, xr_by};
      Prim R{sh_dn(xl_rho), sh_dn(xl_vx), sh_dn(xl_vy), sh_dn(xl_p), s.bxr_b, sh_dn(xl_by)};
      F = hll_x(L, R);
    }

    // ---- Ez(k-1) at the bottom-left corner of cell (c, k-1) ----
    const double fby_left = sh_up(F.b);
    const double gbx_left = sh_up(G.b);
    double ez;
    {
      double eN = -fby_left;
      double eS = -s.fby_left;
      double eE = G.b;
      double eW = gbx_left;
      double vxc = 0.25 * (s.b_vx + l_vx + s.a_vx + s.vx_left_a);
      double vyc = 0.25 * (s.b_vy + l_vy + s.a_vy + s.vy_left_a);
      ez = corner_ez(eN, eS, eE, eW, vxc, vyc);
    }

    // ---- stage-2 resistive term: J(c, k-1) of the stage field ----
    double j_b = 0.0;
    if (MODE == 2) {
      const double byl = sh_up(s.byf_b);
      j_b = a.eta * ((s.byf_b - byl) / a.dx - (s.bxl_b - s.bxl_a) / a.dy);
    }

    // ---- dU(k-2): flux divergence, CT update, stage combination ----
    const int j = k - 2;
    const double fl_rho = sh_up(s.f_rho), fl_mx = sh_up(s.f_mx), fl_my = sh_up(s.f_my), fl_E = sh_up(s.f_E);
    const double ez_right = sh_dn(s.ez);
    const double j_right = (MODE == 2) ? sh_dn(s.j_a) : 0.0;
    if (j >= y0 && j < y1 && out_lane) {
      double d[6];
      d[0] = -((s.f_rho - fl_rho) * idx_ + (G.rho - s.g_rho) * idy_);
      d[1] = -((s.f_mx - fl_mx) * idx_ + (G.mx - s.g_mx) * idy_);
      d[2] = -((s.f_my - fl_my) * idx_ + (G.my - s.g_my) * idy_);
      d[3] = -((s.f_E - fl_E) * idx_ + (G.E - s.g_E) * idy_);
      d[4] = -(ez - s.ez) * idy_;
      d[5] = (ez_right - s.ez) * idx_;
      const int o = j * nx + c0;
      if (MODE == 0) {
#pragma unroll
        for (int q = 0; q < 6; ++q) a.o1[q][o] = d[q];
      } else if (MODE == 1) {
#pragma unroll
        for (int q = 0; q < 6; ++q) {
          a.o1[q][o] = d[q];
          double u = a.s[q][o] + d[q] * a.c_a;
          if (q == 0 && u < kFloor) u = kFloor;
          a.o2[q][o] = u;
        }
      } else {
        double g2x = 0.0, g2y = 0.0;
        if (a.eta != 0.0) {
          g2x = -(j_b - s.j_a) / a.dy;
          g2y = (j_right - s.j_a) / a.dx;
        }
#pragma unroll
        for (int q = 0; q < 6; ++q) {
          double u = a.u[q][o] + a.f1[q][o] * a.c_a;
          u = u + d[q] * a.c_b;
          if (q == 4 && a.eta != 0.0) u = u + g2x * a.c_g;
          if (q == 5 && a.eta != 0.0) u = u + g2y * a.c_g;
          if (q == 0 && u < kFloor) u = kFloor;
          a.o1[q][o] = u;
        }
      }
    }

    // ---- rotate the carried state ----
    s.a_rho = s.b_rho; s.a_vx = s.b_vx; s.a_vy = s.b_vy; s.a_p = s.b_p; s.a_bxc = s.b_bxc;
    s.b_rho = c_rho; s.b_vx = c_vx; s.b_vy = c_vy; s.b_p = c_p; s.b_bxc = bxc; s.b_byc = byc;
    s.yr_rho = yr_rho; s.yr_vx = yr_vx; s.yr_vy = yr_vy; s.yr_p = yr_p; s.yr_bx = yr_bx;
    s.f_rho = F.rho; s.f_mx = F.mx; s.f_my = F.my; s.f_E = F.E;
    s.fby_left = fby_left;
    s.g_rho = G.rho; s.g_mx = G.mx; s.g_my = G.my; s.g_E = G.E;
    s.ez = ez;
    s.vx_left_a = l_vx; s.vy_left_a = l_vy;
    s.bxr_b = bxr;
    s.byf_b = s.byf_c;
    s.byf_c = byn;
    s.j_a = j_b;
    s.bxl_a = s.bxl_b;
    s.bxl_b = bxl;
  }
}

template __global__ void k_stage<0>(StageArgs);
template __global__ void k_stage<1>(StageArgs);
template __global__ void k_stage<2>(StageArgs);

// ---------------------------------------------------------------------------
// Block reductions and the last-block-done pattern.  Partial sums are
// combined in a fixed order, so every reduction is deterministic.
// ---------------------------------------------------------------------------

template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* lds) {
  // wave-level tree (fixed order), then the 4 wave results in order
#pragma unroll
  for (int q = 0; q < NV; ++q) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v[q] += __shfl_down(v[q], off, 64);
  }
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) lds[w * NV + q] = v[q];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = blockDim.x >> 6;
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      double t = lds[q];
      for (int ww = 1; ww < nw; ++ww) t += lds[ww * NV + q];
      v[q] = t;
    }
  }
}

__device__ __forceinline__ double max_ignoring_nan(double m, double v) { return (v > m) ? v : m; }

__device__ __forceinline__ void block_max(double& v, double* lds) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = max_ignoring_nan(v, __shfl_down(v, off, 64));
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) lds[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = blockDim.x >> 6;
    double t = lds[0];
    for (int ww = 1; ww < nw; ++ww) t = max_ignoring_nan(t, lds[ww]);
    v = t;
  }
}

// Returns true in thread 0 of the block that finished last.
__device__ __forceinline__ bool arrive_last(unsigned int* counter, unsigned int nblocks, int* lds_flag) {
  __threadfence();
  if (threadIdx.x == 0) {
    unsigned int prev = atomicAdd(counter, 1u);
    *lds_flag = (prev == nblocks - 1) ? 1 : 0;
  }
  __syncthreads();
  bool last = *lds_flag != 0;
  if (last) __threadfence();
  return last;
}

// Sum NV interleaved partial arrays part[b*NV+q] over nb blocks, fixed order.
template <int NV>
__device__ void final_sum(const double* part, int nb, double (&out)[NV], double* lds) {
  double v[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) v[q] = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
#pragma unroll
    for (int q = 0; q < NV; ++q) v[q] += part[b * NV + q];
  }
  __syncthreads();
  block_sum<NV>(v, lds);
#pragma unroll
  for (int q = 0; q < NV; ++q) out[q] = v[q];
}

// Virtual block index: consecutive workgroup ids are dealt to the 8 XCDs in
// turn; remapping gives each XCD (and its L2) a contiguous band of rows.
__device__ __forceinline__ int xcd_block(int b, int nb) {
  if ((nb & 7) != 0) return b;
  return (b & 7) * (nb >> 3) + (b >> 3);
}

// ---------------------------------------------------------------------------
// Matrix-free Jacobi-preconditioned CG on the corner unknowns, restating
// Eigen's ConjugateGradient (zero initial guess, |r|^2 < max(tol^2|b|^2,
// DBL_MIN), iteration count of completed non-final iterations) on the
// operator of generate_A_matrix (solvers.hpp:404-443):
//   (A x)_ij = d x_ij - cx (x_{i-1,j} + x_{i+1,j}) - cy (x_{i,j-1} + x_{i,j+1}),
//   d = 1 + 2 (cx + cy), cx = alpha / dx^2, cy = alpha / dy^2, periodic.
// The Jacobi preconditioner is the constant 1/d.  Each iteration is two
// passes with one grid-wide reduction each; the stencil product is recomputed
// in the second pass instead of being stored.
// ---------------------------------------------------------------------------

__device__ __forceinline__ void rows_of_block(int vb, int nb, int ny, int& r0, int& r1) {
  int per = (ny + nb - 1) / nb;
  r0 = vb * per;
  r1 = min(r0 + per, ny);
}

// b = eta * J(B) at corners (resistive_B_term's first loop, solvers.hpp:450-463)
// or b = bin; then r = b, p = invd * r; sums |b|^2 and r.p.
__global__ __launch_bounds__(kRedThreads) void k_cg_init(CgArgs a, const double* __restrict__ bin, int explicit_only) {
  __shared__ double lds[16];
  __shared__ int lflag;
  const int nx = a.nx, ny = a.ny;
  const int vb = xcd_block(blockIdx.x, gridDim.x);
  int r0, r1;
  rows_of_block(vb, gridDim.x, ny, r0, r1);
  double acc[2] = {0.0, 0.0};
  for (int j = r0; j < r1; ++j) {
    const int jm = (j == 0) ? ny - 1 : j - 1;
    for (int i = threadIdx.x; i < nx; i += blockDim.x) {
      const int im = (i == 0) ? nx - 1 : i - 1;
      const int o = j * nx + i;
      double b;
      if (bin) {
        b = bin[o];
      } else {
        double dBy_dx = (a.by[o] - a.by[j * nx + im]) / a.dx;
        double dBx_dy = (a.bx[o] - a.bx[jm * nx + i]) / a.dy;
        b = a.eta * (dBy_dx - dBx_dy);
      }
      a.r[o] = b - 0.0;
      if (!explicit_only) {
        double p = a.invd * b;
        a.p0[o] = p;
        acc[0] += b * b;
        acc[1] += b * p;
      }
    }
  }
  if (explicit_only) return;
  block_sum<2>(acc, lds);
  if (threadIdx.x == 0) {
    a.part[vb * 2 + 0] = acc[0];
    a.part[vb * 2 + 1] = acc[1];
  }
  if (!arrive_last(&a.st->counter, gridDim.x, &lflag)) return;
  double tot[2];
  final_sum<2>(a.part, gridDim.x, tot, lds);
  if (threadIdx.x == 0) {
    CgState* st = a.st;
    st->counter = 0;
    st->iters = 0;
    st->spmv = 0;
    st->rhs_norm2 = tot[0];
    st->rn2 = tot[0];
    st->abs_new = tot[1];
    st->alpha = 0.0;
    st->beta = 0.0;
    if (tot[0] == 0.0) {
      st->done = 3;                       // x = 0, iters = 0, error = 0
    } else {
      double thr = dmax(a.tol * a.tol * tot[0], 2.2250738585072014e-308);
      st->threshold = thr;
      if (tot[0] < thr) st->done = 1;     // residual below threshold at x = 0
      else if (st->max_iters <= 0) st->done = 2;
      else st->done = 0;
    }
  }
}

// Pass A: p_new = invd * r + beta * p_old (p_new = p for the first iteration),
// q = A p_new, sum p_new . q -> alpha = absNew / (p . q).
template <bool FIRST>
__global__ __launch_bounds__(kRedThreads) void k_cg_pass_a(CgArgs a, const double* __restrict__ pold,
                                                          double* __restrict__ pnew) {
  __shared__ double lds[16];
  __shared__ int lflag;
  if (a.st->done) return;
  const int nx = a.nx, ny = a.ny;
  const double beta = FIRST ? 0.0 : a.st->beta;
  const double invd = a.invd, d = a.diag, cx = a.cx, cy = a.cy;
  const int vb = xcd_block(blockIdx.x, gridDim.x);
  int r0, r1;
  rows_of_block(vb, gridDim.x, ny, r0, r1);
  double acc[1] = {0.0};
  for (int j = r0; j < r1; ++j) {
    const int jm = (j == 0) ? ny - 1 : j - 1;
    const int jp = (j == ny - 1) ? 0 : j + 1;
    for (int i = threadIdx.x; i < nx; i += blockDim.x) {
      const int im = (i == 0) ? nx - 1 : i - 1;
      const int ip = (i == nx - 1) ? 0 : i + 1;
      const int o = j * nx + i;
      double pc, ps, pw, pe, pn;
      if (FIRST) {
        pc = pold[o]; ps = pold[jm * nx + i]; pw = pold[j * nx + im]; pe = pold[j * nx + ip]; pn = pold[jp * nx + i];
      } else {
        pc = invd * a.r[o] + beta * pold[o];
        ps = invd * a.r[jm * nx + i] + beta * pold[jm * nx + i];
        pw = invd * a.r[j * nx + im] + beta * pold[j * nx + im];
        pe = invd * a.r[j * nx + ip] + beta * pold[j * nx + ip];
        pn = invd * a.r[jp * nx + i] + beta * pold[jp * nx + i];
        pnew[o] = pc;
      }
      double q = -cy * ps;
      q += -cx * pw;
      q += d * pc;
      q += -cx * pe;
      q += -cy * pn;
      acc[0] += pc * q;
    }
  }
  block_sum<1>(acc, lds);
  if (threadIdx.x == 0) a.part[vb] = acc[0];
  if (!arrive_last(&a.st->counter, gridDim.x, &lflag)) return;
  double tot[1];
  final_sum<1>(a.part, gridDim.x, tot, lds);
  if (threadIdx.x == 0) {
    a.st->counter = 0;
    a.st->pap = tot[0];
    a.st->alpha = a.st->abs_new / tot[0];
  }
}

// Pass B: q = A p, x += alpha p, r -= alpha q, sums |r|^2 and r . (invd r);
// convergence test and beta.
template <bool FIRST>
__global__ __launch_bounds__(kRedThreads) void k_cg_pass_b(CgArgs a, const double* __restrict__ pv) {
  __shared__ double lds[16];
  __shared__ int lflag;
  if (a.st->done) return;
  const int nx = a.nx, ny = a.ny;
  const double alpha = a.st->alpha;
  const double invd = a.invd, d = a.diag, cx = a.cx, cy = a.cy;
  const int vb = xcd_block(blockIdx.x, gridDim.x);
  int r0, r1;
  rows_of_block(vb, gridDim.x, ny, r0, r1);
  double acc[2] = {0.0, 0.0};
  for (int j = r0; j < r1; ++j) {
    const int jm = (j == 0) ? ny - 1 : j - 1;
    const int jp = (j == ny - 1) ? 0 : j + 1;
    for (int i = threadIdx.x; i < nx; i += blockDim.x) {
      const int im = (i == 0) ? nx - 1 : i - 1;
      const int ip = (i == nx - 1) ? 0 : i + 1;
      const int o = j * nx + i;
      const double pc = pv[o];
      double q = -cy * pv[jm * nx + i];
      q += -cx * pv[j * nx + im];
      q += d * pc;
      q += -cx * pv[j * nx + ip];
      q += -cy * pv[jp * nx + i];
      const double xo = FIRST ? 0.0 : a.x[o];
      a.x[o] = xo + alpha * pc;
      const double rn = a.r[o] - alpha * q;
      a.r[o] = rn;
      const double z = invd * rn;
      acc[0] += rn * rn;
      acc[1] += rn * z;
    }
  }
  block_sum<2>(acc, lds);
  if (threadIdx.x == 0) {
    a.part[vb * 2 + 0] = acc[0];
    a.part[vb * 2 + 1] = acc[1];
  }
  if (!arrive_last(&a.st->counter, gridDim.x, &lflag)) return;
  double tot[2];
  final_sum<2>(a.part, gridDim.x, tot, lds);
  if (threadIdx.x == 0) {
    CgState* st = a.st;
    st->counter = 0;
    st->spmv += 1;
    st->rn2 = tot[0];
    if (tot[0] < st->threshold) {
      st->done = 1;
    } else {
      double abs_old = st->abs_new;
      st->abs_new = tot[1];
      st->beta = tot[1] / abs_old;
      st->iters += 1;
      if (st->iters >= st->max_iters) st->done = 2;
    }
  }
}

template __global__ void k_cg_pass_a<true>(CgArgs, const double*, double*);
template __global__ void k_cg_pass_a<false>(CgArgs, const double*, double*);
template __global__ void k_cg_pass_b<true>(CgArgs, const double*);
template __global__ void k_cg_pass_b<false>(CgArgs, const double*);

// ---------------------------------------------------------------------------
// Implicit-stage finalisers.
// ---------------------------------------------------------------------------

// rate = -curl_z(x): rate_x(i,j) = -(x(i,j+1) - x(i,j))/dy, rate_y = (x(i+1,j) - x(i,j))/dx
// (resistive_B_term's second loop, solvers.hpp:472-488).
__device__ __forceinline__ double rate_x(const double* x, int nx, int i, int j, int jp, double dy) {
  return -(x[jp * nx + i] - x[j * nx + i]) / dy;
}
__device__ __forceinline__ double rate_y(const double* x, int nx, int i, int j, int ip, double dx) {
  return (x[j * nx + ip] - x[j * nx + i]) / dx;
}

// B += rate(x) * c (solvers.hpp:528-529).  x is taken as 0 when CG did no
// product (Eigen returns x = 0), which leaves B unchanged.
__global__ __launch_bounds__(256) void k_finalize1(FinArgs a) {
  if (a.st->spmv == 0) return;
  const int nx = a.nx, ny = a.ny;
  const int n = nx * ny;
  for (int o = blockIdx.x * blockDim.x + threadIdx.x; o < n; o += gridDim.x * blockDim.x) {
    const int j = o / nx, i = o - j * nx;
    const int jp = (j == ny - 1) ? 0 : j + 1;
    const int ip = (i == nx - 1) ? 0 : i + 1;
    a.bx_out[o] = a.bx[o] + rate_x(a.x, nx, i, j, jp, a.dy) * a.c;
    a.by_out[o] = a.by[o] + rate_y(a.x, nx, i, j, ip, a.dx) * a.c;
  }
}

// Final recombination of the B field (solvers.hpp:568-608): with B3 = B3* +
// c rate(x), B^{n+1} = B3* + c g3, where g3 = -curl(eta J(B3)) is the explicit
// resistive term of B3 (solvers.hpp:579) -- the reference re-evaluates g3
// rather than copying B3.
__device__ __forceinline__ double b3x(const FinArgs& a, bool xv, int i, int j) {
  const int ny = a.ny, nx = a.nx;
  const int jp = (j == ny - 1) ? 0 : j + 1;
  const double r = xv ? rate_x(a.x, nx, i, j, jp, a.dy) : -(0.0 - 0.0) / a.dy;
  return a.bx[j * nx + i] + r * a.c;
}
__device__ __forceinline__ double b3y(const FinArgs& a, bool xv, int i, int j) {
  const int nx = a.nx;
  const int ip = (i == nx - 1) ? 0 : i + 1;
  const double r = xv ? rate_y(a.x, nx, i, j, ip, a.dx) : (0.0 - 0.0) / a.dx;
  return a.by[j * nx + i] + r * a.c;
}
__device__ __forceinline__ double j3(const FinArgs& a, bool xv, int i, int j) {
  const int im = (i == 0) ? a.nx - 1 : i - 1;
  const int jm = (j == 0) ? a.ny - 1 : j - 1;
  const double dBy_dx = (b3y(a, xv, i, j) - b3y(a, xv, im, j)) / a.dx;
  const double dBx_dy = (b3x(a, xv, i, j) - b3x(a, xv, i, jm)) / a.dy;
  return a.eta * (dBy_dx - dBx_dy);
}

__global__ __launch_bounds__(256) void k_finalize2(FinArgs a) {
  const bool xv = a.st->spmv != 0;
  const int nx = a.nx, ny = a.ny;
  const int n = nx * ny;
  for (int o = blockIdx.x * blockDim.x + threadIdx.x; o < n; o += gridDim.x * blockDim.x) {
    const int j = o / nx, i = o - j * nx;
    const int jp = (j == ny - 1) ? 0 : j + 1;
    const int ip = (i == nx - 1) ? 0 : i + 1;
    const double jc = j3(a, xv, i, j);
    const double jt = j3(a, xv, i, jp);
    const double jr = j3(a, xv, ip, j);
    const double g3x = -(jt - jc) / a.dy;
    const double g3y = (jr - jc) / a.dx;
    a.bx_out[o] = a.bx[o] + g3x * a.c;
    a.by_out[o] = a.by[o] + g3y * a.c;
  }
}

// rate arrays of resistive_B_term for the unit entry point.
__global__ __launch_bounds__(256) void k_rate(FinArgs a, int zero) {
  const int nx = a.nx, ny = a.ny;
  const int n = nx * ny;
  for (int o = blockIdx.x * blockDim.x + threadIdx.x; o < n; o += gridDim.x * blockDim.x) {
    const int j = o / nx, i = o - j * nx;
    const int jp = (j == ny - 1) ? 0 : j + 1;
    const int ip = (i == nx - 1) ? 0 : i + 1;
    if (zero) {
      a.bx_out[o] = -(0.0 - 0.0) / a.dy;
      a.by_out[o] = (0.0 - 0.0) / a.dx;
    } else {
      a.bx_out[o] = rate_x(a.x, nx, i, j, jp, a.dy);
      a.by_out[o] = rate_y(a.x, nx, i, j, ip, a.dx);
    }
  }
}

// ---------------------------------------------------------------------------
// Diagnostics and layout helpers.
// ---------------------------------------------------------------------------

// Primitives of one physical cell of the compact state.
__device__ __forceinline__ void cell_prims(const DiagArgs& a, int i, int j, double& rho_f, double& vx, double& vy,
                                           double& p, double& bxc, double& byc) {
  const int nx = a.nx, ny = a.ny;
  const int ip = (i == nx - 1) ? 0 : i + 1;
  const int jp = (j == ny - 1) ? 0 : j + 1;
  const int o = j * nx + i;
  bxc = (a.s[4][o] + a.s[4][j * nx + ip]) / 2.0;
  byc = (a.s[5][o] + a.s[5][jp * nx + i]) / 2.0;
  primitives(a.s[0][o], a.s[1][o], a.s[2][o], a.s[3][o], bxc, byc, rho_f, vx, vy, p);
}

// CFL_condition_checker's max wave speed (diagnostics.hpp:58-80); NaN speeds
// are skipped exactly as the reference's `>` comparison skips them, and
// counted separately.
__global__ __launch_bounds__(kRedThreads) void k_cfl(DiagArgs a) {
  __shared__ double lds[16];
  __shared__ int lflag;
  const int nx = a.nx, ny = a.ny;
  const int vb = xcd_block(blockIdx.x, gridDim.x);
  int r0, r1;
  rows_of_block(vb, gridDim.x, ny, r0, r1);
  double m = 1e-8;
  double bad = 0.0;
  for (int j = r0; j < r1; ++j) {
    for (int i = threadIdx.x; i < nx; i += blockDim.x) {
      double rho, vx, vy, p, bx, by;
      cell_prims(a, i, j, rho, vx, vy, p, bx, by);
      rho = dmax(rho, 1e-8);
      p = dmax(p, 1e-8);
      double ss = kGamma * p / rho;
      double as = (bx * bx + by * by) / rho;
      double ms = sqrt(ss + as);
      double fs = sqrt(vx * vx + vy * vy);
      double cur = ms + fs;
      if (cur > m) m = cur;
      if (!(fabs(cur) <= 1.7976931348623157e308)) bad += 1.0;
    }
  }
  double b1[1] = {bad};
  block_max(m, lds);
  __syncthreads();
  block_sum<1>(b1, lds + 8);
  if (threadIdx.x == 0) {
    a.part[vb * 2 + 0] = m;
    a.part[vb * 2 + 1] = b1[0];
  }
  if (!arrive_last(a.counter, gridDim.x, &lflag)) return;
  double mm = 1e-8, nb = 0.0;
  for (int b = threadIdx.x; b < (int)gridDim.x; b += blockDim.x) {
    mm = max_ignoring_nan(mm, a.part[b * 2]);
    nb += a.part[b * 2 + 1];
  }
  __syncthreads();
  block_max(mm, lds);
  __syncthreads();
  double nb1[1] = {nb};
  block_sum<1>(nb1, lds + 8);
  if (threadIdx.x == 0) {
    *a.counter = 0;
    a.out[0] = mm;
    a.out[1] = nb1[0];
  }
}

// calculate_current_div_B (diagnostics.hpp:15-42): sum of squares and max |div|.
__global__ __launch_bounds__(kRedThreads) void k_divb(DiagArgs a) {
  __shared__ double lds[16];
  __shared__ int lflag;
  const int nx = a.nx, ny = a.ny;
  const int vb = xcd_block(blockIdx.x, gridDim.x);
  int r0, r1;
  rows_of_block(vb, gridDim.x, ny, r0, r1);
  double ss[1] = {0.0};
  double m = 0.0;
  for (int j = r0; j < r1; ++j) {
    const int jp = (j == ny - 1) ? 0 : j + 1;
    for (int i = threadIdx.x; i < nx; i += blockDim.x) {
      const int ip = (i == nx - 1) ? 0 : i + 1;
      const int o = j * nx + i;
      double xd = (a.s[4][j * nx + ip] - a.s[4][o]) / a.dx;
      double yd = (a.s[5][jp * nx + i] - a.s[5][o]) / a.dy;
      double t = xd + yd;
      ss[0] += t * t;
      if (fabs(t) > m) m = fabs(t);
    }
  }
  block_sum<1>(ss, lds);
  __syncthreads();
  block_max(m, lds + 8);
  if (threadIdx.x == 0) {
    a.part[vb * 2 + 0] = ss[0];
    a.part[vb * 2 + 1] = m;
  }
  if (!arrive_last(a.counter, gridDim.x, &lflag)) return;
  double s1[1] = {0.0};
  double mm = 0.0;
  for (int b = threadIdx.x; b < (int)gridDim.x; b += blockDim.x) {
    s1[0] += a.part[b * 2];
    mm = max_ignoring_nan(mm, a.part[b * 2 + 1]);
  }
  __syncthreads();
  block_sum<1>(s1, lds);
  __syncthreads();
  block_max(mm, lds + 8);
  if (threadIdx.x == 0) {
    *a.counter = 0;
    a.out[0] = s1[0];
    a.out[1] = mm;
  }
}

// Primitives of every cell, for mhd_download.
__global__ __launch_bounds__(256) void k_prims(DiagArgs a, double* vx, double* vy, double* p, double* bxc,
                                               double* byc) {
  const int n = a.nx * a.ny;
  for (int o = blockIdx.x * blockDim.x + threadIdx.x; o < n; o += gridDim.x * blockDim.x) {
    const int j = o / a.nx, i = o - j * a.nx;
    double r, u, v, pp, bx, by;
    cell_prims(a, i, j, r, u, v, pp, bx, by);
    vx[o] = u; vy[o] = v; p[o] = pp; bxc[o] = bx; byc[o] = by;
  }
}

// Density floor applied to a freshly uploaded state (Grid.hpp:115).
__global__ __launch_bounds__(256) void k_floor_rho(double* rho, int n) {
  for (int o = blockIdx.x * blockDim.x + threadIdx.x; o < n; o += gridDim.x * blockDim.x)
    if (rho[o] < kFloor) rho[o] = kFloor;
}

__global__ __launch_bounds__(256) void k_copy(const double* __restrict__ src, double* __restrict__ dst, long long n) {
  for (long long o = blockIdx.x * (long long)blockDim.x + threadIdx.x; o < n; o += (long long)gridDim.x * blockDim.x)
    dst[o] = src[o];
}

}  // namespace mhd
