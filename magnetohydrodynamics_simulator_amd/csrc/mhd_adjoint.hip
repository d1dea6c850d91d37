// mhd_adjoint.hip -- discrete adjoint of the ARS(2,2,2) IMEX step (C ABI
// mhd_adjoint_*), with checkpointing.
//
// The step of mhd_imex_step, for a state U = (h, B) (h: rho, mx, my, E at
// cells; B: face bx, by), constant dt, c = dt gamma:
//   f1 = L(U);  W = U + dt delta f1;  S* = U + dt gamma f1 (rho floored)
//   x1 = A^-1 eta J(S*.B);  S = (S*.h, S*.B + c R(x1))
//   U3* = W + dt (1-delta) L(S) + dt (1-gamma) R(eta J(S.B)) [on B] (rho floored)
//   x2 = A^-1 eta J(U3*.B);  B3 = U3*.B + c R(x2)
//   U' = (U3*.h, U3*.B + c R(eta J(B3)))
// with L the explicit MUSCL-HLL-CT right-hand side, J the corner current
// (solvers.hpp:450-463), R the face rate -curl_z (solvers.hpp:472-488) and
// A = I - alpha Lap, alpha = dt eta gamma, symmetric.  The adjoint runs these
// lines backwards: transposes of J, R and the flux differences, the adjoint
// of L by reverse-mode differentiation of primitives, MUSCL-minmod, HLL and
// the upwind CT selection at the recorded forward state (branches taken as in
// the forward pass; the CT upwind choice and the minmod selection are
// piecewise constant), the density/pressure floors as 0/1 masks, and the
// adjoint of each implicit solve as a second CG solve with the same
// symmetric A.  d/d eta collects the eta of J (three places) and of alpha.
//
// Driver: mhd_adjoint_forward runs n steps of constant dt from the current
// state keeping a copy of the state every k steps; mhd_adjoint_backward
// takes the adjoint of the final state, and for each k-step segment from the
// last restores the checkpoint, recomputes the segment recording every
// step's stage data, and runs the adjoint steps backwards.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mhd_mi355x.h"
#include "mhd_context.hpp"
#include "mhd_adjoint_math.hpp"
#include "mhd_device.hpp"
#include "mhd_kernels.hpp"

using namespace mhd;
using namespace mhd_host;

namespace mhd {

// ---------------------------------------------------------------------------
// Linear stencils of the resistive terms and their transposes (periodic).
//   J(B)(i,j)  = (by(i,j) - by(i-1,j))/dx - (bx(i,j) - bx(i,j-1))/dy   (corner)
//   R(x)       = (-(x(i,j+1) - x(i,j))/dy, (x(i+1,j) - x(i,j))/dx)    (faces)
//   Lap(x)     = (x(i-1,j) - 2x + x(i+1,j))/dx^2 + (x(i,j-1) - 2x + x(i,j+1))/dy^2
// ---------------------------------------------------------------------------

struct LinArgs {
  int nx, ny;
  double idx, idy;
};

// q = J(bx, by)
__global__ __launch_bounds__(256) void ka_J(LinArgs g, const double* bx, const double* by, double* q) {
  const int n = g.nx * g.ny;
  for (int o = blockIdx.x * blockDim.x + threadIdx.x; o < n; o += gridDim.x * blockDim.x) {
    const int j = o / g.nx, i = o - j * g.nx;
    q[o] = (by[o] - by[j * g.nx + wp(i - 1, g.nx)]) * g.idx - (bx[o] - bx[wp(j - 1, g.ny) * g.nx + i]) * g.idy;
  }
}

// (fx, fy) += c J^T(q)
__global__ __launch_bounds__(256) void ka_Jt(LinArgs g, const double* q, double c, double* fx, double* fy) {
  const int n = g.nx * g.ny;
  for (int o = blockIdx.x * blockDim.x + threadIdx.x; o < n; o += gridDim.x * blockDim.x) {
    const int j = o / g.nx, i = o - j * g.nx;
    fx[o] += c * (q[wp(j + 1, g.ny) * g.nx + i] - q[o]) * g.idy;
    fy[o] += c * (q[o] - q[j * g.nx + wp(i + 1, g.nx)]) * g.idx;
  }
}

// (fx, fy) = (bx, by) + c R(x)
__global__ __launch_bounds__(256) void ka_R(LinArgs g, const double* bx, const double* by, const double* x, double c,
                                            double* fx, double* fy) {
  const int n = g.nx * g.ny;
  for (int o = blockIdx.x * blockDim.x + threadIdx.x; o < n; o += gridDim.x * blockDim.x) {
    const int j = o / g.nx, i = o - j * g.nx;
    const double xc = x ? x[o] : 0.0;
    const double xn = x ? x[wp(j + 1, g.ny) * g.nx + i] : 0.0;
    const double xe = x ? x[j * g.nx + wp(i + 1, g.nx)] : 0.0;
    fx[o] = bx[o] + c * (-(xn - xc) * g.idy);
    fy[o] = by[o] + c * ((xe - xc) * g.idx);
  }
}

// q = c R^T(fx, fy)
__global__ __launch_bounds__(256) void ka_Rt(LinArgs g, const double* fx, const double* fy, double c, double* q) {
  const int n = g.nx * g.ny;
  for (int o = blockIdx.x * blockDim.x + threadIdx.x; o < n; o += gridDim.x * blockDim.x) {
    const int j = o / g.nx, i = o - j * g.nx;
    q[o] = c * ((fx[o] - fx[wp(j - 1, g.ny) * g.nx + i]) * g.idy + (fy[j * g.nx + wp(i - 1, g.nx)] - fy[o]) * g.idx);
  }
}

// partial[b] = sum over the block's cells of w . (a or Lap a)
__global__ __launch_bounds__(256) void ka_dot(LinArgs g, const double* u, const double* v, int lap, double* partial) {
  __shared__ double lds[16];
  const int n = g.nx * g.ny;
  double acc[1] = {0.0};
  for (int o = blockIdx.x * blockDim.x + threadIdx.x; o < n; o += gridDim.x * blockDim.x) {
    double vv = v[o];
    if (lap) {
      const int j = o / g.nx, i = o - j * g.nx;
      vv = (v[j * g.nx + wp(i - 1, g.nx)] - 2.0 * v[o] + v[j * g.nx + wp(i + 1, g.nx)]) * (g.idx * g.idx) +
           (v[wp(j - 1, g.ny) * g.nx + i] - 2.0 * v[o] + v[wp(j + 1, g.ny) * g.nx + i]) * (g.idy * g.idy);
    }
    acc[0] += u[o] * vv;
  }
  // fixed-order block reduction
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc[0] += __shfl_down(acc[0], off, 64);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) lds[w] = acc[0];
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = lds[0];
    for (int ww = 1; ww < (int)(blockDim.x >> 6); ++ww) t += lds[ww];
    partial[blockIdx.x] = t;
  }
}

// acc[0] += scale * sum(partial[0..nb)) in a fixed order (one thread).
__global__ void ka_accumulate(const double* partial, int nb, double scale, double* acc) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    double t = 0.0;
    for (int b = 0; b < nb; ++b) t += partial[b];
    acc[0] += scale * t;
  }
}

// y = a x + b y on nf fields
__global__ __launch_bounds__(256) void ka_axpby(int n, double a, const double* x, double b, double* y) {
  for (int o = blockIdx.x * blockDim.x + threadIdx.x; o < n; o += gridDim.x * blockDim.x) y[o] = a * x[o] + b * y[o];
}

// y = a x + b z (elementwise, fresh output)
__global__ __launch_bounds__(256) void ka_lin2(int n, double a, const double* x, double b, const double* z, double* y) {
  for (int o = blockIdx.x * blockDim.x + threadIdx.x; o < n; o += gridDim.x * blockDim.x) y[o] = a * x[o] + b * z[o];
}

// y *= (ref > floor): adjoint of the density floor u = max(u, 1e-8)
__global__ __launch_bounds__(256) void ka_floor_mask(int n, const double* ref, double* y) {
  for (int o = blockIdx.x * blockDim.x + threadIdx.x; o < n; o += gridDim.x * blockDim.x)
    if (!(ref[o] > kFloor)) y[o] = 0.0;
}

__global__ __launch_bounds__(256) void ka_zero(int n, double* y) {
  for (int o = blockIdx.x * blockDim.x + threadIdx.x; o < n; o += gridDim.x * blockDim.x) y[o] = 0.0;
}

}  // namespace mhd

// ---------------------------------------------------------------------------
// Host side.
// ---------------------------------------------------------------------------

namespace mhd_host {

struct AdjointStore {
  double dt = 0.0;
  int nsteps = 0, every = 0;
  std::vector<double*> ckpt;          // 6 fields per checkpoint (steps 0, k, 2k, ...)
  std::vector<StepRecord> rec;        // one per step of a segment
  std::vector<double*> owned;         // every device buffer of this store
  double* P[7] = {};
  double* xrb[5] = {};
  double* xlb[5] = {};
  double* yrb[5] = {};
  double* ylb[5] = {};
  double* bnx = nullptr;
  double* bny = nullptr;
  double* bcb[2] = {};
  double* lam[6] = {};                // adjoint of the state after the step (in), before it (out)
  double* wb[6] = {};                 // W_bar = U3*_bar
  double* sb[6] = {};                 // S_bar
  double* fb[6] = {};                 // f_bar (adjoint of an explicit rate)
  double* q1 = nullptr;               // corner scratch
  double* q2 = nullptr;
  double* fx = nullptr;               // face scratch
  double* fy = nullptr;
  double* partial = nullptr;
  double* eta_bar = nullptr;          // device accumulator
  int nblk = 0;
};

void destroy_adjoint(mhd_ctx* c) {
  if (!c->adj) return;
  for (double* p : c->adj->owned)
    if (p) (void)hipFree(p);
  delete c->adj;
  c->adj = nullptr;
}

}  // namespace mhd_host

namespace {

int alloc(mhd_ctx* c, double** p) {
  if (hipMalloc(p, (size_t)c->n * sizeof(double)) != hipSuccess) {
    *p = nullptr;
    return fail(c, MHD_ENOMEM, "adjoint: device allocation failed");
  }
  c->adj->owned.push_back(*p);
  return MHD_OK;
}

int ensure_store(mhd_ctx* c, int seg) {
  if (!c->adj) {
    c->adj = new AdjointStore();
    AdjointStore& A = *c->adj;
    double** fields[] = {A.P, A.xrb, A.xlb, A.yrb, A.ylb, A.bcb, A.lam, A.wb, A.sb, A.fb};
    const int counts[] = {7, 5, 5, 5, 5, 2, 6, 6, 6, 6};
    for (int g = 0; g < 10; ++g)
      for (int q = 0; q < counts[g]; ++q)
        if (int rc = alloc(c, &fields[g][q])) return rc;
    double** singles[] = {&A.bnx, &A.bny, &A.q1, &A.q2, &A.fx, &A.fy};
    for (double** p : singles)
      if (int rc = alloc(c, p)) return rc;
    A.nblk = (int)std::min<long long>((c->n + 255) / 256, 1024);
    if (hipMalloc(&A.partial, sizeof(double) * A.nblk) != hipSuccess) return fail(c, MHD_ENOMEM, "adjoint: allocation failed");
    A.owned.push_back(A.partial);
    if (hipMalloc(&A.eta_bar, sizeof(double)) != hipSuccess) return fail(c, MHD_ENOMEM, "adjoint: allocation failed");
    A.owned.push_back(A.eta_bar);
  }
  AdjointStore& A = *c->adj;
  while ((int)A.rec.size() < seg) {
    StepRecord r{};
    for (int q = 0; q < 6; ++q) {
      if (int rc = alloc(c, &r.Un[q])) return rc;
      if (int rc = alloc(c, &r.S[q])) return rc;
    }
    for (int q = 0; q < 2; ++q) {
      if (int rc = alloc(c, &r.SsB[q])) return rc;
      if (int rc = alloc(c, &r.U3sB[q])) return rc;
    }
    if (int rc = alloc(c, &r.x1)) return rc;
    if (int rc = alloc(c, &r.x2)) return rc;
    A.rec.push_back(r);
  }
  return MHD_OK;
}

LinArgs lin_args(mhd_ctx* c) {
  LinArgs g;
  g.nx = c->nx;
  g.ny = c->ny;
  g.idx = 1.0 / c->dx;
  g.idy = 1.0 / c->dy;
  return g;
}

// out += L'(s)^T lam for the state s (6 device fields).
int explicit_rhs_vjp(mhd_ctx* c, double* const* s, double* const* lam, double* const* out) {
  AdjointStore& A = *c->adj;
  AdjArgs a;
  a.nx = c->nx;
  a.ny = c->ny;
  a.dx = c->dx;
  a.dy = c->dy;
  a.idx = 1.0 / c->dx;
  a.idy = 1.0 / c->dy;
  for (int q = 0; q < 6; ++q) {
    a.s[q] = s[q];
    a.lam[q] = lam[q];
    a.out[q] = out[q];
  }
  for (int q = 0; q < 7; ++q) a.P[q] = A.P[q];
  for (int q = 0; q < 5; ++q) {
    a.xrb[q] = A.xrb[q];
    a.xlb[q] = A.xlb[q];
    a.yrb[q] = A.yrb[q];
    a.ylb[q] = A.ylb[q];
  }
  a.bnx = A.bnx;
  a.bny = A.bny;
  a.bcb[0] = A.bcb[0];
  a.bcb[1] = A.bcb[1];
  const int g = grid_1d(c->n);
  int rc = launch(c, 8, [&] { hipLaunchKernelGGL(ka_prims, dim3(g), dim3(256), 0, c->stream, a); });
  if (rc) return rc;
  rc = launch(c, 8, [&] { hipLaunchKernelGGL(ka_faces, dim3(g), dim3(256), 0, c->stream, a); });
  if (rc) return rc;
  rc = launch(c, 8, [&] { hipLaunchKernelGGL(ka_cells, dim3(g), dim3(256), 0, c->stream, a); });
  if (rc) return rc;
  return launch(c, 8, [&] { hipLaunchKernelGGL(ka_bfaces, dim3(g), dim3(256), 0, c->stream, a); });
}

// eta_bar += scale * <u, v or Lap v>
int dot_acc(mhd_ctx* c, const double* u, const double* v, bool lap, double scale) {
  AdjointStore& A = *c->adj;
  LinArgs g = lin_args(c);
  int rc = launch(c, 8, [&] {
    hipLaunchKernelGGL(ka_dot, dim3(A.nblk), dim3(256), 0, c->stream, g, u, v, lap ? 1 : 0, A.partial);
  });
  if (rc) return rc;
  return launch(c, 8, [&] {
    hipLaunchKernelGGL(ka_accumulate, dim3(1), dim3(1), 0, c->stream, (const double*)A.partial, A.nblk, scale, A.eta_bar);
  });
}

// Adjoint of one implicit solve x = A^-1 eta J(B) (B the face field the
// solve's right-hand side was formed from, x the forward solution):
// given xbar (corner field, in q), adds eta J^T A^-1 xbar to (bxb, byb) and
// the solve's eta derivative <J(B), bbar> + dt gamma <bbar, Lap x> to eta_bar.
int solve_vjp(mhd_ctx* c, double* xbar, const double* const* B, const double* x, bool have_x, double* bxb,
              double* byb) {
  AdjointStore& A = *c->adj;
  LinArgs g = lin_args(c);
  const int gr = grid_1d(c->n);
  CgState cs;
  int rc = cg_run(c, nullptr, nullptr, xbar, &cs, 0);   // bbar = A^-1 xbar in c->x
  if (rc) return rc;
  const double* bbar = c->x;
  if (cs.spmv == 0) {                                     // zero right-hand side: bbar = 0
    rc = launch(c, 8, [&] { hipLaunchKernelGGL(ka_zero, dim3(gr), dim3(256), 0, c->stream, (int)c->n, c->x); });
    if (rc) return rc;
  }
  rc = launch(c, 8, [&] { hipLaunchKernelGGL(ka_Jt, dim3(gr), dim3(256), 0, c->stream, g, bbar, c->eta, bxb, byb); });
  if (rc) return rc;
  rc = launch(c, 8, [&] { hipLaunchKernelGGL(ka_J, dim3(gr), dim3(256), 0, c->stream, g, B[0], B[1], A.q2); });
  if (rc) return rc;
  rc = dot_acc(c, A.q2, bbar, false, 1.0);
  if (rc) return rc;
  if (have_x) {
    rc = dot_acc(c, bbar, x, true, A.dt * c->tab.A_im4);
    if (rc) return rc;
  }
  return MHD_OK;
}

// Adjoint of the explicit resistive term g = R(eta J(B)) for gbar (faces):
// adds eta J^T R^T gbar to (bxb, byb) and <J(B), R^T gbar> to eta_bar.
int resistive_vjp(mhd_ctx* c, const double* gx, const double* gy, double coef, const double* const* B, double* bxb,
                  double* byb) {
  AdjointStore& A = *c->adj;
  LinArgs g = lin_args(c);
  const int gr = grid_1d(c->n);
  int rc = launch(c, 8, [&] { hipLaunchKernelGGL(ka_Rt, dim3(gr), dim3(256), 0, c->stream, g, gx, gy, coef, A.q1); });
  if (rc) return rc;
  rc = launch(c, 8, [&] { hipLaunchKernelGGL(ka_Jt, dim3(gr), dim3(256), 0, c->stream, g, (const double*)A.q1, c->eta, bxb, byb); });
  if (rc) return rc;
  rc = launch(c, 8, [&] { hipLaunchKernelGGL(ka_J, dim3(gr), dim3(256), 0, c->stream, g, B[0], B[1], A.q2); });
  if (rc) return rc;
  return dot_acc(c, A.q2, A.q1, false, 1.0);
}

// lam (adjoint of the state after the step) -> adjoint of the state before it.
// rec: the step's forward data; rho_next: density after the step.
int adjoint_step(mhd_ctx* c, const StepRecord& rec, const double* rho_next) {
  AdjointStore& A = *c->adj;
  const Tableau& T = c->tab;
  const double dt = A.dt, cg = dt * T.A_im4;
  const int n = (int)c->n;
  const int gr = grid_1d(c->n);
  LinArgs g = lin_args(c);
  int rc;
  double** lam = A.lam;
  // ---- final recombination: U' = (U3*.h, U3*.B + cg R(eta J(B3))), B3 = U3*.B + cg R(x2)
  // B3 into (fx, fy)
  rc = launch(c, 8, [&] {
    hipLaunchKernelGGL(ka_R, dim3(gr), dim3(256), 0, c->stream, g, (const double*)rec.U3sB[0], (const double*)rec.U3sB[1],
                       rec.spmv[1] > 0 ? (const double*)rec.x2 : nullptr, cg, A.fx, A.fy);
  });
  if (rc) return rc;
  // B3bar = eta J^T R^T (cg lam.B) into wb.B (zeroed first); eta_bar += <J(B3), R^T(cg lam.B)>
  for (int q = 4; q < 6; ++q) {
    rc = launch(c, 8, [&] { hipLaunchKernelGGL(ka_zero, dim3(gr), dim3(256), 0, c->stream, n, A.wb[q]); });
    if (rc) return rc;
  }
  {
    const double* B3[2] = {A.fx, A.fy};
    rc = resistive_vjp(c, lam[4], lam[5], cg, B3, A.wb[4], A.wb[5]);
    if (rc) return rc;
  }
  // x2bar = cg R^T(B3bar) -> solve adjoint: U3*.Bbar += eta J^T A^-1 x2bar
  rc = launch(c, 8, [&] { hipLaunchKernelGGL(ka_Rt, dim3(gr), dim3(256), 0, c->stream, g, (const double*)A.wb[4], (const double*)A.wb[5], cg, A.q1); });
  if (rc) return rc;
  // U3*.Bbar = lam.B + B3bar (+ solve term below), accumulated in wb.B
  for (int q = 4; q < 6; ++q) {
    rc = launch(c, 8, [&] { hipLaunchKernelGGL(ka_axpby, dim3(gr), dim3(256), 0, c->stream, n, 1.0, (const double*)lam[q], 1.0, A.wb[q]); });
    if (rc) return rc;
  }
  {
    const double* B[2] = {rec.U3sB[0], rec.U3sB[1]};
    rc = solve_vjp(c, A.q1, B, rec.x2, rec.spmv[1] > 0, A.wb[4], A.wb[5]);
    if (rc) return rc;
  }
  // hydro: U3*.hbar = lam.h; density floor of U3*
  for (int q = 0; q < 4; ++q) {
    rc = launch(c, 8, [&] { hipLaunchKernelGGL(ka_lin2, dim3(gr), dim3(256), 0, c->stream, n, 1.0, (const double*)lam[q], 0.0, (const double*)lam[q], A.wb[q]); });
    if (rc) return rc;
  }
  rc = launch(c, 8, [&] { hipLaunchKernelGGL(ka_floor_mask, dim3(gr), dim3(256), 0, c->stream, n, rho_next, A.wb[0]); });
  if (rc) return rc;
  // ---- stage 2: U3* = W + dt(1-delta) L(S) + dt(1-gamma) R(eta J(S.B)); Wbar = U3*bar (wb)
  // Sbar = eta J^T R^T (dt(1-gamma) U3*.Bbar) + L'(S)^T (dt(1-delta) U3*bar)
  for (int q = 0; q < 6; ++q) {
    rc = launch(c, 8, [&] { hipLaunchKernelGGL(ka_zero, dim3(gr), dim3(256), 0, c->stream, n, A.sb[q]); });
    if (rc) return rc;
    rc = launch(c, 8, [&] { hipLaunchKernelGGL(ka_lin2, dim3(gr), dim3(256), 0, c->stream, n, dt * T.A_ex7, (const double*)A.wb[q], 0.0, (const double*)A.wb[q], A.fb[q]); });
    if (rc) return rc;
  }
  {
    const double* SB[2] = {rec.S[4], rec.S[5]};
    rc = resistive_vjp(c, A.wb[4], A.wb[5], dt * T.A_im7, SB, A.sb[4], A.sb[5]);
    if (rc) return rc;
  }
  rc = explicit_rhs_vjp(c, rec.S, A.fb, A.sb);
  if (rc) return rc;
  // ---- implicit 1: S.B = S*.B + cg R(x1), x1 = A^-1 eta J(S*.B): S*bar = Sbar + solve term
  rc = launch(c, 8, [&] { hipLaunchKernelGGL(ka_Rt, dim3(gr), dim3(256), 0, c->stream, g, (const double*)A.sb[4], (const double*)A.sb[5], cg, A.q1); });
  if (rc) return rc;
  {
    const double* B[2] = {rec.SsB[0], rec.SsB[1]};
    rc = solve_vjp(c, A.q1, B, rec.x1, rec.spmv[0] > 0, A.sb[4], A.sb[5]);
    if (rc) return rc;
  }
  // density floor of S* (S.rho = S*.rho)
  rc = launch(c, 8, [&] { hipLaunchKernelGGL(ka_floor_mask, dim3(gr), dim3(256), 0, c->stream, n, (const double*)rec.S[0], A.sb[0]); });
  if (rc) return rc;
  // ---- stage 1: S* = U + dt gamma f1, W = U + dt delta f1, f1 = L(U)
  // Ubar = S*bar + Wbar + L'(U)^T (dt gamma S*bar + dt delta Wbar)
  for (int q = 0; q < 6; ++q) {
    rc = launch(c, 8, [&] { hipLaunchKernelGGL(ka_lin2, dim3(gr), dim3(256), 0, c->stream, n, dt * T.A_ex3, (const double*)A.sb[q], dt * T.A_ex6, (const double*)A.wb[q], A.fb[q]); });
    if (rc) return rc;
    rc = launch(c, 8, [&] { hipLaunchKernelGGL(ka_lin2, dim3(gr), dim3(256), 0, c->stream, n, 1.0, (const double*)A.sb[q], 1.0, (const double*)A.wb[q], lam[q]); });
    if (rc) return rc;
  }
  return explicit_rhs_vjp(c, rec.Un, A.fb, lam);
}

int check(mhd_ctx* c) { return c ? MHD_OK : MHD_EINVAL; }

}  // namespace

extern "C" {

int mhd_adjoint_forward(mhd_ctx* c, double dt, int nsteps, int checkpoint_every) {
  if (check(c)) return MHD_EINVAL;
  if (!(dt > 0) || !std::isfinite(dt) || nsteps < 1 || checkpoint_every < 1)
    return fail(c, MHD_EINVAL, "mhd_adjoint_forward: need dt > 0, nsteps >= 1, checkpoint_every >= 1");
  if (!(c->eta > 0)) return fail(c, MHD_EINVAL, "mhd_adjoint_forward: the adjoint needs eta > 0");
  int rc = ensure_store(c, std::min(checkpoint_every, nsteps));
  if (rc) return rc;
  AdjointStore& A = *c->adj;
  const int nck = (nsteps + checkpoint_every - 1) / checkpoint_every;
  while ((int)A.ckpt.size() < 6 * nck) {
    double* p = nullptr;
    if ((rc = alloc(c, &p))) return rc;
    A.ckpt.push_back(p);
  }
  A.dt = dt;
  A.nsteps = nsteps;
  A.every = checkpoint_every;
  rc = mhd_build_operator(c, (dt * c->eta) * c->tab.A_im4);   // constant dt: one operator
  if (rc) return rc;
  const size_t fb = (size_t)c->n * sizeof(double);
  for (int s = 0; s < nsteps; ++s) {
    if (s % checkpoint_every == 0)
      for (int q = 0; q < 6; ++q)
        HIP_OR_FAIL(c, hipMemcpyAsync(A.ckpt[6 * (s / checkpoint_every) + q], c->U[q], fb, hipMemcpyDeviceToDevice, c->stream));
    rc = imex_step(c, dt, nullptr, nullptr);
    if (rc) return rc;
  }
  HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
  return MHD_OK;
}

int mhd_adjoint_backward(mhd_ctx* c, const double* const* lam_final, double* const* lam0, double* eta_bar) {
  if (check(c)) return MHD_EINVAL;
  if (!c->adj || c->adj->nsteps < 1) return fail(c, MHD_EINVAL, "mhd_adjoint_backward before mhd_adjoint_forward");
  if (!lam_final) return fail(c, MHD_EINVAL, "lam_final is NULL");
  for (int q = 0; q < 6; ++q)
    if (!lam_final[q]) return fail(c, MHD_EINVAL, "lam_final has a NULL field");
  AdjointStore& A = *c->adj;
  const size_t fb = (size_t)c->n * sizeof(double);
  int rc;
  for (int q = 0; q < 6; ++q) {
    rc = from_host_padded(c, lam_final[q], A.lam[q]);
    if (rc) return rc;
  }
  const double zero = 0.0;
  HIP_OR_FAIL(c, hipMemcpyAsync(A.eta_bar, &zero, sizeof(double), hipMemcpyHostToDevice, c->stream));
  // the final state (density after the last step) is the current state
  double* final_rho = nullptr;
  if ((rc = alloc(c, &final_rho))) return rc;
  HIP_OR_FAIL(c, hipMemcpyAsync(final_rho, c->U[0], fb, hipMemcpyDeviceToDevice, c->stream));
  // keep the current (final) state: the recompute overwrites U
  double* keep[6];
  for (int q = 0; q < 6; ++q) {
    if ((rc = alloc(c, &keep[q]))) return rc;
    HIP_OR_FAIL(c, hipMemcpyAsync(keep[q], c->U[q], fb, hipMemcpyDeviceToDevice, c->stream));
  }
  const int nseg = (A.nsteps + A.every - 1) / A.every;
  const double* rho_next = final_rho;
  for (int sg = nseg - 1; sg >= 0; --sg) {
    const int s0 = sg * A.every, s1 = std::min(A.nsteps, s0 + A.every);
    for (int q = 0; q < 6; ++q)
      HIP_OR_FAIL(c, hipMemcpyAsync(c->U[q], A.ckpt[6 * sg + q], fb, hipMemcpyDeviceToDevice, c->stream));
    for (int s = s0; s < s1; ++s) {
      rc = imex_step(c, A.dt, nullptr, &A.rec[s - s0]);
      if (rc) return rc;
    }
    for (int s = s1 - 1; s >= s0; --s) {
      const StepRecord& r = A.rec[s - s0];
      const double* rn = (s == A.nsteps - 1) ? final_rho : (s + 1 < s1 ? A.rec[s + 1 - s0].Un[0] : rho_next);
      rc = adjoint_step(c, r, rn);
      if (rc) return rc;
    }
    rho_next = A.ckpt[6 * sg + 0];
  }
  for (int q = 0; q < 6; ++q)
    HIP_OR_FAIL(c, hipMemcpyAsync(c->U[q], keep[q], fb, hipMemcpyDeviceToDevice, c->stream));
  if (lam0)
    for (int q = 0; q < 6; ++q)
      if (lam0[q]) {
        rc = to_host_padded(c, A.lam[q], lam0[q], true);
        if (rc) return rc;
      }
  double eb = 0.0;
  HIP_OR_FAIL(c, hipMemcpyAsync(&eb, A.eta_bar, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
  if (eta_bar) *eta_bar = eb;
  // release the temporaries
  auto& own = A.owned;
  for (double* p : keep) {
    own.erase(std::remove(own.begin(), own.end(), p), own.end());
    (void)hipFree(p);
  }
  own.erase(std::remove(own.begin(), own.end(), final_rho), own.end());
  (void)hipFree(final_rho);
  return MHD_OK;
}

int mhd_explicit_rhs_vjp(mhd_ctx* c, const double* const* lam, double* const* out) {
  if (check(c)) return MHD_EINVAL;
  if (!lam || !out) return fail(c, MHD_EINVAL, "mhd_explicit_rhs_vjp: NULL argument");
  int rc = ensure_store(c, 0);
  if (rc) return rc;
  AdjointStore& A = *c->adj;
  const int gr = grid_1d(c->n);
  for (int q = 0; q < 6; ++q) {
    if (!lam[q]) return fail(c, MHD_EINVAL, "mhd_explicit_rhs_vjp: NULL field");
    rc = from_host_padded(c, lam[q], A.fb[q]);
    if (rc) return rc;
    rc = launch(c, 8, [&] { hipLaunchKernelGGL(ka_zero, dim3(gr), dim3(256), 0, c->stream, (int)c->n, A.sb[q]); });
    if (rc) return rc;
  }
  rc = explicit_rhs_vjp(c, c->U, A.fb, A.sb);
  if (rc) return rc;
  for (int q = 0; q < 6; ++q)
    if (out[q]) {
      rc = to_host_padded(c, A.sb[q], out[q], false);
      if (rc) return rc;
    }
  return MHD_OK;
}

}  // extern "C"
