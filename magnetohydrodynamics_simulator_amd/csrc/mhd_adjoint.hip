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
#include "mhd_adjoint_lin.hpp"
#include "mhd_adjoint_march.hpp"
#include "mhd_adjoint_math.hpp"
#include "mhd_device.hpp"
#include "mhd_frozen.hpp"
#include "mhd_kernels.hpp"

using namespace mhd;
using namespace mhd_host;

namespace mhd {

// Single-run kernels: grid-stride loops over the cells of one grid running the
// bodies of mhd_adjoint_lin.hpp.
#define FOR_CELLS(g)                                                                        \
  for (int o_ = blockIdx.x * blockDim.x + threadIdx.x, n_ = (g).nx * (g).ny; o_ < n_;       \
       o_ += gridDim.x * blockDim.x)                                                        \
    if (const int j_ = o_ / (g).nx, i_ = o_ - j_ * (g).nx; true)

// Adjoint of the explicit right-hand side, one kernel per body of
// mhd_adjoint_math.hpp.
__global__ __launch_bounds__(256) void ka_prims(AdjArgs a) {
  const int n = a.nx * a.ny;
  for (int o = blockIdx.x * blockDim.x + threadIdx.x; o < n; o += gridDim.x * blockDim.x) {
    const int j = o / a.nx;
    ka_prims_cell(a, o - j * a.nx, j);
  }
}

__global__ __launch_bounds__(256) void ka_faces(AdjArgs a) {
  const int n = a.nx * a.ny;
  for (int o = blockIdx.x * blockDim.x + threadIdx.x; o < n; o += gridDim.x * blockDim.x) {
    const int j = o / a.nx;
    ka_faces_cell(a, o - j * a.nx, j);
  }
}

__global__ __launch_bounds__(256) void ka_cells(AdjArgs a) {
  const int n = a.nx * a.ny;
  for (int o = blockIdx.x * blockDim.x + threadIdx.x; o < n; o += gridDim.x * blockDim.x) {
    const int j = o / a.nx;
    ka_cells_cell(a, o - j * a.nx, j);
  }
}

__global__ __launch_bounds__(256) void ka_bfaces(AdjArgs a) {
  const int n = a.nx * a.ny;
  for (int o = blockIdx.x * blockDim.x + threadIdx.x; o < n; o += gridDim.x * blockDim.x) {
    const int j = o / a.nx;
    ka_bfaces_cell(a, o - j * a.nx, j);
  }
}

// The fused adjoint of the explicit right-hand side (mhd_adjoint_march.hpp):
// one wavefront per (strip, segment).
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(MHD_MARCH_WAVES))) void ka_march(MarchArgs a) { adj_march(a, blockIdx.x, blockIdx.y); }

__global__ __launch_bounds__(256) void ka_Jt(LinArgs g, const double* q, double c, double* fx, double* fy) {
  FOR_CELLS(g) Jt_cell(g, q, c, fx, fy, i_, j_);
}

__global__ __launch_bounds__(256) void ka_Jt2(LinArgs g, const double* q, double c, const double* bx, const double* by,
                                              double* tx, double* ty, double* ox, double* oy) {
  FOR_CELLS(g) Jt2_cell(g, q, c, bx, by, tx, ty, ox, oy, i_, j_);
}

__global__ __launch_bounds__(256) void ka_R(LinArgs g, const double* bx, const double* by, const double* x,
                                            const CgState* st, double c, double* fx, double* fy) {
  const bool zx = x_is_zero(st);
  FOR_CELLS(g) R_cell(g, bx, by, x, zx, c, fx, fy, i_, j_);
}

__global__ __launch_bounds__(256) void ka_Rt(LinArgs g, const double* fx, const double* fy, double c, double* q) {
  FOR_CELLS(g) Rt_cell(g, fx, fy, c, q, i_, j_);
}

// partial[b] = block sum of u . Lap(v)  (0 if v is a solution of a solve with no product)
__global__ __launch_bounds__(256) void ka_dot_lap(LinArgs g, const double* u, const double* v, const CgState* st,
                                                  double* partial) {
  double acc = 0.0;
  if (!x_is_zero(st)) FOR_CELLS(g) acc += lap_dot_cell(g, u, v, i_, j_);
  block_partial(acc, partial + blockIdx.x);
}

// partial[b] = block sum of q . J(bx, by)
__global__ __launch_bounds__(256) void ka_dot_J(LinArgs g, const double* q, const double* bx, const double* by,
                                                double* partial) {
  double acc = 0.0;
  FOR_CELLS(g) acc += q[o_] * corner_J(g, bx, by, i_, j_);
  block_partial(acc, partial + blockIdx.x);
}

struct EtaScales {
  double s[6];
};

// eta_bar += sum_k scale_k * sum_b partial[k * nb + b], in a fixed order.
__global__ __launch_bounds__(256) void ka_accumulate(const double* partial, int nslots, int nb, EtaScales sc,
                                                     double* eta_bar) {
  __shared__ double lds[8];
  double e = 0.0;
  for (int k = 0; k < nslots; ++k) {
    double acc = 0.0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) acc += partial[k * nb + b];
    e += sc.s[k] * block_total(acc, lds);
  }
  if (threadIdx.x == 0) eta_bar[0] += e;
}

// x = 0 where the solve did no product
__global__ __launch_bounds__(256) void ka_zero_if_no_product(const CgState* st, int n, double* x) {
  if (st->spmv != 0) return;
  for (int o = blockIdx.x * blockDim.x + threadIdx.x; o < n; o += gridDim.x * blockDim.x) x[o] = 0.0;
}

__global__ __launch_bounds__(256) void ka_adj_stage2_in(int n, Fields6 lam, Fields6 wb, Fields6 fb, Fields6 sb,
                                                        const double* rho_next, double ca) {
  for (int o = blockIdx.x * blockDim.x + threadIdx.x; o < n; o += gridDim.x * blockDim.x)
    stage2_in_cell(o, lam, wb, fb, sb, rho_next, ca);
}

__global__ __launch_bounds__(256) void ka_adj_stage1_out(int n, Fields6 sb, Fields6 wb, Fields6 fb, Fields6 lam,
                                                         const double* rho_s, double cs, double cw) {
  for (int o = blockIdx.x * blockDim.x + threadIdx.x; o < n; o += gridDim.x * blockDim.x)
    stage1_out_cell(o, sb, wb, fb, lam, rho_s, cs, cw);
}

// one block per Jacobi tile, tiles numbered x fastest
__global__ __launch_bounds__(256) void ka_jacobi(LinArgs g, double cx, double cy, double invd, const double* b,
                                                 const double* xin, double* xout) {
  const int tiles_x = (g.nx + kJacTile - 1) / kJacTile;
  jacobi_tile<kJacTile, kJacSweeps>(g, cx, cy, invd, b, xin, xout, blockIdx.x % tiles_x, blockIdx.x / tiles_x);
}

__global__ __launch_bounds__(256) void ka_zero(int n, double* y) {
  for (int o = blockIdx.x * blockDim.x + threadIdx.x; o < n; o += gridDim.x * blockDim.x) y[o] = 0.0;
}

// Branch words of a stage state whose primitives ka_prims has just written.
__global__ __launch_bounds__(256) void ka_record(AdjArgs a, const double* rho_after, uint64_t* out) {
  FOR_CELLS(a) out[o_] = record_cell(a, rho_after, i_, j_);
}

// The explicit right-hand side with recorded branch choices (mhd_frozen.hpp).
__global__ __launch_bounds__(256) void kf_prims(FrozenArgs a) { FOR_CELLS(a) fz_prims_cell(a, i_, j_); }
__global__ __launch_bounds__(256) void kf_faces(FrozenArgs a) { FOR_CELLS(a) fz_faces_cell(a, i_, j_); }
__global__ __launch_bounds__(256) void kf_update(FrozenArgs a) { FOR_CELLS(a) fz_update_cell(a, i_, j_); }

__global__ __launch_bounds__(256) void kf_comb1(int n, Fields6 u, Fields6 f, double c1, double cw, const uint64_t* br,
                                                Fields6 s, Fields6 w) {
  for (int o = blockIdx.x * blockDim.x + threadIdx.x; o < n; o += gridDim.x * blockDim.x)
    fz_comb1_cell(o, u.f, f.f, c1, cw, br, s.f, w.f);
}

__global__ __launch_bounds__(256) void kf_comb2(LinArgs g, Fields6 w, Fields6 f, const double* q, double c2,
                                                double c3, const uint64_t* br, Fields6 u) {
  FOR_CELLS(g) fz_comb2_cell(g, i_, j_, w.f, f.f, q, c2, c3, br, u.f);
}

// q = eta J(bx, by)
__global__ __launch_bounds__(256) void kf_J(LinArgs g, const double* bx, const double* by, double eta, double* q) {
  FOR_CELLS(g) q[o_] = eta * corner_J(g, bx, by, i_, j_);
}

#undef FOR_CELLS

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
  std::vector<int> fwd_spmv;          // operator products of each forward solve (2 per step)
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
  double* keep[6] = {};               // the final state, kept across the recompute
  double* final_rho = nullptr;
  double* q1 = nullptr;               // corner scratch
  double* fx = nullptr;               // face scratch
  double* fy = nullptr;
  double* gx = nullptr;
  double* gy = nullptr;
  double* partial = nullptr;          // 6 slots of nblk block sums (eta derivative terms)
  double* eta_bar = nullptr;          // device accumulator
  std::vector<double*> xs;            // the two implicit solutions of every forward step
  CgState* fwd_states = nullptr;      // their final solver states (2 per step)
  double* jac = nullptr;              // Jacobi ping-pong buffer
  // branch words of every stage of the last recorded backward pass
  // (2 per step, N each: mhd_adjoint_record_branches / mhd_adjoint_forward_frozen)
  uint64_t* br = nullptr;
  long long br_cap = 0;               // words allocated
  bool record = false;                // record in the next backward pass
  bool use_graph = true;              // replay the sync-free backward sweep as a captured graph
  int br_steps = 0;                   // steps recorded (0: none)
  double br_dt = 0.0;
  CgState* states = nullptr;          // final states of every solve of a backward pass (4 per step)
  CgState* states_host = nullptr;     // pinned mirror
  int nstates = 0;
  int nblk = 0;
  // cached graph of a sync-free backward sweep and what it was captured for
  hipGraphExec_t graph = nullptr;
  double* g_entry[12] = {};           // U and S arrays at capture
  double* g_exit[12] = {};            // and after the sweep
  int g_nsteps = 0, g_every = 0;
  double g_dt = 0.0, g_alpha = 0.0;
  const void* g_xs = nullptr;
};

void destroy_adjoint(mhd_ctx* c) {
  if (!c->adj) return;
  // the graph's nodes name the buffers below: destroy it first
  if (c->adj->graph) (void)hipGraphExecDestroy(c->adj->graph);
  c->adj->graph = nullptr;
  for (double* p : c->adj->owned)
    if (p) (void)hipFree(p);
  if (c->adj->states) (void)hipFree(c->adj->states);
  if (c->adj->fwd_states) (void)hipFree(c->adj->fwd_states);
  if (c->adj->br) (void)hipFree(c->adj->br);
  if (c->adj->states_host) (void)hipHostFree(c->adj->states_host);
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

int ensure_store(mhd_ctx* c, int seg, int nsteps) {
  if (!c->adj) {
    c->adj = new AdjointStore();
    AdjointStore& A = *c->adj;
    double** fields[] = {A.P, A.xrb, A.xlb, A.yrb, A.ylb, A.bcb, A.lam, A.wb, A.sb, A.fb, A.keep};
    const int counts[] = {7, 5, 5, 5, 5, 2, 6, 6, 6, 6, 6};
    for (int g = 0; g < 11; ++g)
      for (int q = 0; q < counts[g]; ++q)
        if (int rc = alloc(c, &fields[g][q])) return rc;
    double** singles[] = {&A.bnx, &A.bny, &A.q1, &A.fx, &A.fy, &A.gx, &A.gy, &A.final_rho, &A.jac};
    for (double** p : singles)
      if (int rc = alloc(c, p)) return rc;
    A.nblk = (int)std::min<long long>((c->n + 255) / 256, 1024);
    if (hipMalloc(&A.partial, sizeof(double) * 6 * A.nblk) != hipSuccess) return fail(c, MHD_ENOMEM, "adjoint: allocation failed");
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
  if (A.nstates < 4 * nsteps) {
    if (A.states) (void)hipFree(A.states);
    if (A.states_host) (void)hipHostFree(A.states_host);
    if (A.fwd_states) (void)hipFree(A.fwd_states);
    A.states = nullptr;
    A.states_host = nullptr;
    A.fwd_states = nullptr;
    A.nstates = 0;
    if (hipMalloc(&A.states, sizeof(CgState) * 4 * nsteps) != hipSuccess ||
        hipHostMalloc(&A.states_host, sizeof(CgState) * 4 * nsteps) != hipSuccess ||
        hipMalloc(&A.fwd_states, sizeof(CgState) * 2 * nsteps) != hipSuccess)
      return fail(c, MHD_ENOMEM, "adjoint: allocation failed");
    A.nstates = 4 * nsteps;
  }
  while ((int)A.xs.size() < 2 * nsteps) {
    double* p = nullptr;
    if (int rc = alloc(c, &p)) return rc;
    A.xs.push_back(p);
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

Fields6 f6(double* const* p) {
  Fields6 f;
  for (int q = 0; q < 6; ++q) f.f[q] = p[q];
  return f;
}

// out += L'(s)^T lam for the state s (6 device fields), by the four unfused
// gather kernels (the reference of the fused kernel).
AdjArgs adj_args(mhd_ctx* c, double* const* s, double* const* lam, double* const* out) {
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
    a.lam[q] = lam ? lam[q] : nullptr;
    a.out[q] = out ? out[q] : nullptr;
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
  return a;
}

int explicit_rhs_vjp_unfused(mhd_ctx* c, double* const* s, double* const* lam, double* const* out) {
  const AdjArgs a = adj_args(c, s, lam, out);
  const int g = grid_1d(c->n);
  int rc = launch(c, 8, [&] { hipLaunchKernelGGL(ka_prims, dim3(g), dim3(256), 0, c->stream, a); });
  if (rc) return rc;
  rc = launch(c, 8, [&] { hipLaunchKernelGGL(ka_faces, dim3(g), dim3(256), 0, c->stream, a); });
  if (rc) return rc;
  rc = launch(c, 8, [&] { hipLaunchKernelGGL(ka_cells, dim3(g), dim3(256), 0, c->stream, a); });
  if (rc) return rc;
  return launch(c, 8, [&] { hipLaunchKernelGGL(ka_bfaces, dim3(g), dim3(256), 0, c->stream, a); });
}

// Rows per segment of the fused adjoint march: 64, fewer when that leaves
// fewer than about 2048 wavefronts (8 per CU), at least 8.
int march_rows(int nx, int ny) {
  const long long strips = (nx + kAdjOut - 1) / kAdjOut;
  long long rows = 64;
  while (rows > 8 && strips * ((ny + rows - 1) / rows) < 2048) rows /= 2;
  return (int)rows;
}

// Grids below this many cells take the unfused kernels: the fused march is
// a chain of (rows + 7) dependent row steps per wavefront, about 75 us
// whatever the size, while the four unfused kernels are fully parallel
// (measured on MI355X: unfused faster up to 512^2, fused from 768^2 on,
// 1.4x at 1024^2 and 1.9x at 2048^2).
constexpr long long kFusedMinCells = 1LL << 19;

// out += L'(s)^T lam for the state s (6 device fields), by the fused march.
int explicit_rhs_vjp_fused(mhd_ctx* c, double* const* s, double* const* lam, double* const* out) {
  MarchArgs a;
  for (int q = 0; q < 6; ++q) {
    a.s[q] = s[q];
    a.lam[q] = lam[q];
    a.out[q] = out[q];
  }
  a.nx = c->nx;
  a.ny = c->ny;
  a.rows = march_rows(c->nx, c->ny);
  a.idx = 1.0 / c->dx;
  a.idy = 1.0 / c->dy;
  const dim3 grid((c->nx + kAdjOut - 1) / kAdjOut, (c->ny + a.rows - 1) / a.rows);
  return launch(c, 8, [&] { hipLaunchKernelGGL(ka_march, grid, dim3(64), 0, c->stream, a); });
}

// out += L'(s)^T lam, by whichever of the two is faster at this size.  With
// words != nullptr also the branch words of s (the unfused kernels then run,
// so that the words come from the primitives the product used); rho_after:
// the density of the state the stage produces.
int explicit_rhs_vjp(mhd_ctx* c, double* const* s, double* const* lam, double* const* out,
                     uint64_t* words = nullptr, const double* rho_after = nullptr) {
  if (words) {
    int rc = explicit_rhs_vjp_unfused(c, s, lam, out);
    if (rc) return rc;
    const AdjArgs a = adj_args(c, s, nullptr, nullptr);
    return launch(c, 8, [&] {
      hipLaunchKernelGGL(ka_record, dim3(grid_1d(c->n)), dim3(256), 0, c->stream, a, rho_after, words);
    });
  }
  if (c->n < kFusedMinCells) return explicit_rhs_vjp_unfused(c, s, lam, out);
  return explicit_rhs_vjp_fused(c, s, lam, out);
}

// Jacobi sweeps of the adjoint solves with this context's operator (0: CG).
int jacobi_sweeps(mhd_ctx* c) {
  return jacobi_sweep_count(c->alpha / (c->dx * c->dx), c->alpha / (c->dy * c->dy));
}

// bbar = A^-1 xbar in c->x.  A is symmetric, so the adjoint of a solve is a
// solve with the same operator: by Jacobi sweeps in fused launches of 8 when
// the operator is close enough to the identity (see jacobi_sweeps), else by
// CG (zero where the solve does no product; its final state is copied to
// *state; spec > 0: speculative, see cg_run).
// c->x = A^-1 b by `sweeps` Jacobi sweeps from x = 0 (rounded up to whole
// launches of kJacSweeps) for A = I - alpha Lap with cx = alpha / dx^2,
// cy = alpha / dy^2.
int jacobi_run(mhd_ctx* c, double cx, double cy, int sweeps, const double* b) {
  AdjointStore& A = *c->adj;
  const double invd = 1.0 / (1 + 2 * (cx + cy));
  LinArgs g = lin_args(c);
  const int tiles = ((c->nx + kJacTile - 1) / kJacTile) * ((c->ny + kJacTile - 1) / kJacTile);
  const int launches = (sweeps + kJacSweeps - 1) / kJacSweeps;
  // ping-pong so that the last launch writes c->x
  double* bufs[2] = {(launches & 1) ? c->x : A.jac, (launches & 1) ? A.jac : c->x};
  const double* in = nullptr;
  for (int k = 0; k < launches; ++k) {
    double* out = bufs[k & 1];
    int rc = launch(c, 8, [&] {
      hipLaunchKernelGGL(ka_jacobi, dim3(tiles), dim3(256), 0, c->stream, g, cx, cy, invd, b, in, out);
    });
    if (rc) return rc;
    in = out;
  }
  return MHD_OK;
}

int adjoint_solve(mhd_ctx* c, const double* xbar, int spec, CgState* state) {
  const int sweeps = jacobi_sweeps(c);
  if (sweeps > 0) return jacobi_run(c, c->alpha / (c->dx * c->dx), c->alpha / (c->dy * c->dy), sweeps, xbar);
  CgState cs;
  int rc = cg_run(c, nullptr, nullptr, xbar, &cs, 2, spec, state);
  if (rc) return rc;
  if (spec == 0) HIP_OR_FAIL(c, hipMemcpyAsync(state, c->st, sizeof(CgState), hipMemcpyDeviceToDevice, c->stream));
  const int gr = grid_1d(c->n);
  return launch(c, 8, [&] {
    hipLaunchKernelGGL(ka_zero_if_no_product, dim3(gr), dim3(256), 0, c->stream, (const CgState*)c->st, (int)c->n, c->x);
  });
}

// lam (adjoint of the state after the step) -> adjoint of the state before it.
// rec: the step's forward data; rho_next: density after the step; spec:
// speculative iterations of the two adjoint solves (0: host-checked);
// states: where their final states go.
// step: the step's index (for the branch words when they are recorded).
int adjoint_step(mhd_ctx* c, const StepRecord& rec, const double* rho_next, int spec, CgState* states, int step) {
  AdjointStore& A = *c->adj;
  const Tableau& T = c->tab;
  const double dt = A.dt, cg = dt * T.A_im4, eta = c->eta;
  const int n = (int)c->n;
  const int gr = grid_1d(c->n);
  const int nb = A.nblk;
  LinArgs g = lin_args(c);
  double** lam = A.lam;
  double* P = A.partial;
  int rc;
#define KL(...) do { rc = launch(c, 8, [&] { hipLaunchKernelGGL(__VA_ARGS__); }); if (rc) return rc; } while (0)
  // ---- final recombination: U' = (U3*.h, U3*.B + cg R(eta J(B3))), B3 = U3*.B + cg R(x2)
  KL(ka_R, dim3(gr), dim3(256), 0, c->stream, g, (const double*)rec.U3sB[0], (const double*)rec.U3sB[1],
     (const double*)rec.x2, (const CgState*)rec.st[1], cg, A.fx, A.fy);
  // g3bar = cg lam.B: q1 = R^T g3bar; B3bar = eta J^T q1 (gx, gy); wb.B = lam.B + B3bar
  KL(ka_Rt, dim3(gr), dim3(256), 0, c->stream, g, (const double*)lam[4], (const double*)lam[5], cg, A.q1);
  KL(ka_Jt2, dim3(gr), dim3(256), 0, c->stream, g, (const double*)A.q1, eta, (const double*)lam[4],
     (const double*)lam[5], A.gx, A.gy, A.wb[4], A.wb[5]);
  KL(ka_dot_J, dim3(nb), dim3(256), 0, c->stream, g, (const double*)A.q1, (const double*)A.fx, (const double*)A.fy, P + 0 * nb);
  // x2bar = cg R^T B3bar; b2bar = A^-1 x2bar; U3*.Bbar += eta J^T b2bar
  KL(ka_Rt, dim3(gr), dim3(256), 0, c->stream, g, (const double*)A.gx, (const double*)A.gy, cg, A.q1);
  rc = adjoint_solve(c, A.q1, spec, states + 1);
  if (rc) return rc;
  KL(ka_Jt, dim3(gr), dim3(256), 0, c->stream, g, (const double*)c->x, eta, A.wb[4], A.wb[5]);
  KL(ka_dot_J, dim3(nb), dim3(256), 0, c->stream, g, (const double*)c->x, (const double*)rec.U3sB[0],
     (const double*)rec.U3sB[1], P + 1 * nb);
  KL(ka_dot_lap, dim3(nb), dim3(256), 0, c->stream, g, (const double*)c->x, (const double*)rec.x2,
     (const CgState*)rec.st[1], P + 2 * nb);
  // ---- stage 2: U3* = W + dt(1-delta) L(S) + dt(1-gamma) R(eta J(S.B)); Wbar = U3*bar
  KL(ka_adj_stage2_in, dim3(gr), dim3(256), 0, c->stream, n, f6(lam), f6(A.wb), f6(A.fb), f6(A.sb), rho_next,
     dt * T.A_ex7);
  KL(ka_Rt, dim3(gr), dim3(256), 0, c->stream, g, (const double*)A.wb[4], (const double*)A.wb[5], dt * T.A_im7, A.q1);
  KL(ka_Jt, dim3(gr), dim3(256), 0, c->stream, g, (const double*)A.q1, eta, A.sb[4], A.sb[5]);
  KL(ka_dot_J, dim3(nb), dim3(256), 0, c->stream, g, (const double*)A.q1, (const double*)rec.S[4],
     (const double*)rec.S[5], P + 3 * nb);
  uint64_t* words2 = A.record ? A.br + (2LL * step + 1) * c->n : nullptr;
  rc = explicit_rhs_vjp(c, rec.S, A.fb, A.sb, words2, rho_next);
  if (rc) return rc;
  // ---- implicit 1: S.B = S*.B + cg R(x1), x1 = A^-1 eta J(S*.B)
  KL(ka_Rt, dim3(gr), dim3(256), 0, c->stream, g, (const double*)A.sb[4], (const double*)A.sb[5], cg, A.q1);
  rc = adjoint_solve(c, A.q1, spec, states + 0);
  if (rc) return rc;
  KL(ka_Jt, dim3(gr), dim3(256), 0, c->stream, g, (const double*)c->x, eta, A.sb[4], A.sb[5]);
  KL(ka_dot_J, dim3(nb), dim3(256), 0, c->stream, g, (const double*)c->x, (const double*)rec.SsB[0],
     (const double*)rec.SsB[1], P + 4 * nb);
  KL(ka_dot_lap, dim3(nb), dim3(256), 0, c->stream, g, (const double*)c->x, (const double*)rec.x1,
     (const CgState*)rec.st[0], P + 5 * nb);
  // ---- stage 1: S* = U + dt gamma f1, W = U + dt delta f1, f1 = L(U)
  KL(ka_adj_stage1_out, dim3(gr), dim3(256), 0, c->stream, n, f6(A.sb), f6(A.wb), f6(A.fb), f6(lam),
     (const double*)rec.S[0], dt * T.A_ex3, dt * T.A_ex6);
  uint64_t* words1 = A.record ? A.br + 2LL * step * c->n : nullptr;
  rc = explicit_rhs_vjp(c, rec.Un, A.fb, lam, words1, rec.S[0]);
  if (rc) return rc;
  // ---- d/d eta: <J(B3), R^T g3bar> + <J(U3*.B), b2bar> + dt gamma <b2bar, Lap x2>
  //               + <J(S.B), R^T g2bar> + <J(S*.B), b1bar> + dt gamma <b1bar, Lap x1>
  EtaScales sc{{1.0, 1.0, dt * T.A_im4, 1.0, 1.0, dt * T.A_im4}};
  KL(ka_accumulate, dim3(1), dim3(256), 0, c->stream, (const double*)P, 6, nb, sc, A.eta_bar);
#undef KL
  return MHD_OK;
}

int check(mhd_ctx* c) { return c ? MHD_OK : MHD_EINVAL; }

// The adjoint kernels wrap periodically: outflow contexts are refused.
int check_periodic(mhd_ctx* c, const char* what) {
  if (c->o.bc_x != MHD_BC_PERIODIC || c->o.bc_y != MHD_BC_PERIODIC)
    return fail(c, MHD_EINVAL, std::string(what) + ": the adjoint is periodic (bc_x = bc_y = MHD_BC_PERIODIC)");
  return MHD_OK;
}

// The backward sweep; spec: run every solve speculatively (verified by the caller).
int backward(mhd_ctx* c, bool spec) {
  AdjointStore& A = *c->adj;
  const size_t fb = (size_t)c->n * sizeof(double);

  const int nseg = (A.nsteps + A.every - 1) / A.every;
  const double* rho_next = A.final_rho;
  for (int sg = nseg - 1; sg >= 0; --sg) {
    const int s0 = sg * A.every, s1 = std::min(A.nsteps, s0 + A.every);
    int rc0 = copy_fields(c, c->U, &A.ckpt[6 * sg], 6);
    if (rc0) return rc0;
    for (int s = s0; s < s1; ++s) {
      // the recompute takes the implicit solutions the forward pass stored
      StepRecord& r = A.rec[s - s0];
      r.use_given = true;
      r.x1 = A.xs[2 * s];
      r.x2 = A.xs[2 * s + 1];
      r.st[0] = A.fwd_states + 2 * s;
      r.st[1] = A.fwd_states + 2 * s + 1;
      int rc = imex_step(c, A.dt, nullptr, &r);
      if (rc) return rc;
    }
    for (int s = s1 - 1; s >= s0; --s) {
      const StepRecord& r = A.rec[s - s0];
      const double* rn = (s == A.nsteps - 1) ? A.final_rho : (s + 1 < s1 ? A.rec[s + 1 - s0].Un[0] : rho_next);
      // the adjoint solves use the operator of the forward ones; their
      // right-hand sides differ, so a margin is added to the forward counts
      const int sp = std::max(A.fwd_spmv[2 * s], A.fwd_spmv[2 * s + 1]);
      const int spec_adj = spec ? sp + 2 + sp / 4 : 0;
      int rc = adjoint_step(c, r, rn, spec_adj, A.states + 4 * s + 2, s);
      if (rc) return rc;
    }
    rho_next = A.ckpt[6 * sg + 0];
  }
  return MHD_OK;
}

}  // namespace

namespace {

int rhs_vjp_entry(mhd_ctx* c, const double* const* lam, double* const* out, bool fused) {
  if (check(c)) return MHD_EINVAL;
  if (!lam || !out) return fail(c, MHD_EINVAL, "mhd_explicit_rhs_vjp: NULL argument");
  if (int rcp = check_periodic(c, "mhd_explicit_rhs_vjp")) return rcp;
  int rc = ensure_store(c, 0, 0);
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
  rc = fused ? explicit_rhs_vjp_fused(c, c->U, A.fb, A.sb) : explicit_rhs_vjp_unfused(c, c->U, A.fb, A.sb);
  if (rc) return rc;
  for (int q = 0; q < 6; ++q)
    if (out[q]) {
      rc = to_host_padded(c, A.sb[q], out[q], false);
      if (rc) return rc;
    }
  return MHD_OK;
}

}  // namespace

extern "C" {

int mhd_adjoint_forward(mhd_ctx* c, double dt, int nsteps, int checkpoint_every) {
  if (check(c)) return MHD_EINVAL;
  if (int rcp = check_periodic(c, "mhd_adjoint_forward")) return rcp;
  if (!(dt > 0) || !std::isfinite(dt) || nsteps < 1 || checkpoint_every < 1)
    return fail(c, MHD_EINVAL, "mhd_adjoint_forward: need dt > 0, nsteps >= 1, checkpoint_every >= 1");
  if (!(c->eta > 0)) return fail(c, MHD_EINVAL, "mhd_adjoint_forward: the adjoint needs eta > 0");
  int rc = ensure_store(c, std::min(checkpoint_every, nsteps), nsteps);
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
  A.fwd_spmv.assign(2 * nsteps, 1);
  rc = mhd_build_operator(c, (dt * c->eta) * c->tab.A_im4);   // constant dt: one operator
  if (rc) return rc;
  for (int s = 0; s < nsteps; ++s) {
    if (s % checkpoint_every == 0) {
      rc = copy_fields(c, &A.ckpt[6 * (s / checkpoint_every)], c->U, 6);
      if (rc) return rc;
    }
    mhd_step_stats st;
    StepRecord r{};
    r.x1 = A.xs[2 * s];
    r.x2 = A.xs[2 * s + 1];
    r.st[0] = A.fwd_states + 2 * s;
    r.st[1] = A.fwd_states + 2 * s + 1;
    rc = imex_step(c, dt, &st, &r);
    if (rc) return rc;
    A.fwd_spmv[2 * s] = st.cg_spmv[0];
    A.fwd_spmv[2 * s + 1] = st.cg_spmv[1];
  }
  HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
  return MHD_OK;
}

}  // extern "C"

namespace {

// The sweep of mhd_adjoint_backward between the upload of lam_final and the
// download of the results: keep the final state, zero eta_bar, run the
// backward pass, restore the final state.
int sweep(mhd_ctx* c, bool spec) {
  AdjointStore& A = *c->adj;
  int rc = copy_fields(c, A.keep, c->U, 6);
  if (rc) return rc;
  rc = copy_fields(c, &A.final_rho, c->U, 1);
  if (rc) return rc;
  rc = launch(c, 8, [&] { hipLaunchKernelGGL(ka_zero, dim3(1), dim3(64), 0, c->stream, 1, A.eta_bar); });
  if (rc) return rc;
  rc = backward(c, spec);
  if (rc) return rc;
  rc = copy_fields(c, c->U, A.keep, 6);
  if (rc) return rc;
  // the recompute overwrote the unfloored densities: the next CFL step reads
  // the final state's primitives
  return reset_raw_density(c);
}

// mhd_adjoint_backward with the final adjoint read from, and the initial one
// written to, padded host arrays (device == false) or compact device arrays.
int backward_entry(mhd_ctx* c, const double* const* lam_final, double* const* lam0, double* eta_bar, bool device) {
  if (check(c)) return MHD_EINVAL;
  if (!c->adj || c->adj->nsteps < 1) return fail(c, MHD_EINVAL, "mhd_adjoint_backward before mhd_adjoint_forward");
  if (!lam_final) return fail(c, MHD_EINVAL, "lam_final is NULL");
  for (int q = 0; q < 6; ++q)
    if (!lam_final[q]) return fail(c, MHD_EINVAL, "lam_final has a NULL field");
  AdjointStore& A = *c->adj;
  int rc;
  auto upload_lam = [&]() -> int {
    if (device) {
      double* src[6];
      for (int q = 0; q < 6; ++q) src[q] = const_cast<double*>(lam_final[q]);
      return copy_fields(c, A.lam, src, 6);
    }
    for (int q = 0; q < 6; ++q) {
      int r = from_host_padded(c, lam_final[q], A.lam[q]);
      if (r) return r;
    }
    return MHD_OK;
  };
  if (A.record) {
    // the record is filled by this pass: allocate it for 2 words per step and cell
    const long long need = 2LL * A.nsteps * c->n;
    if (A.br_cap < need) {
      if (A.br) (void)hipFree(A.br);
      A.br = nullptr;
      A.br_cap = 0;
      if (hipMalloc(&A.br, sizeof(uint64_t) * need) != hipSuccess) {
        A.br = nullptr;
        return fail(c, MHD_ENOMEM, "mhd_adjoint_backward: branch record allocation failed");
      }
      A.br_cap = need;
    }
    A.br_steps = 0;
  }
  if (jacobi_sweeps(c) > 0 && c->prof_mask == 0 && !A.record && A.use_graph) {
    // no host round trip anywhere in the sweep: replay it as one graph
    // (captured on the first call of a configuration, reused after)
    rc = upload_lam();
    if (rc) return rc;
    bool same = A.graph && A.g_nsteps == A.nsteps && A.g_every == A.every && A.g_dt == A.dt && A.g_alpha == c->alpha &&
                A.g_xs == (const void*)A.xs.data();
    for (int q = 0; q < 6 && same; ++q) same = A.g_entry[q] == c->U[q] && A.g_entry[6 + q] == c->S[q];
    if (!same) {
      if (A.graph) (void)hipGraphExecDestroy(A.graph);
      A.graph = nullptr;
      for (int q = 0; q < 6; ++q) {
        A.g_entry[q] = c->U[q];
        A.g_entry[6 + q] = c->S[q];
      }
      HIP_OR_FAIL(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      rc = sweep(c, false);
      hipGraph_t g = nullptr;
      hipError_t e = hipStreamEndCapture(c->stream, &g);
      if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
      if (e != hipSuccess) return fail(c, MHD_EHIP, std::string("graph capture: ") + hipGetErrorString(e));
      e = hipGraphInstantiate(&A.graph, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (e != hipSuccess) {
        A.graph = nullptr;
        return fail(c, MHD_EHIP, std::string("graph instantiate: ") + hipGetErrorString(e));
      }
      for (int q = 0; q < 6; ++q) {
        A.g_exit[q] = c->U[q];
        A.g_exit[6 + q] = c->S[q];
      }
      A.g_nsteps = A.nsteps;
      A.g_every = A.every;
      A.g_dt = A.dt;
      A.g_alpha = c->alpha;
      A.g_xs = (const void*)A.xs.data();
    }
    HIP_OR_FAIL(c, hipGraphLaunch(A.graph, c->stream));
    for (int q = 0; q < 6; ++q) {
      c->U[q] = A.g_exit[q];
      c->S[q] = A.g_exit[6 + q];
    }
  } else {
    // first pass: every adjoint CG solve speculative (no host round trip),
    // verified afterwards; a solve that did not finish sends the sweep round
    // again with host-checked solves
    for (int attempt = 0; attempt < 2; ++attempt) {
      const bool spec = (attempt == 0);
      rc = upload_lam();
      if (rc) return rc;
      rc = sweep(c, spec);
      if (rc) return rc;
      HIP_OR_FAIL(c, hipMemcpyAsync(A.states_host, A.states, sizeof(CgState) * 4 * A.nsteps, hipMemcpyDeviceToHost,
                                    c->stream));
      HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
      // only CG solves leave a state to check (Jacobi sweeps always finish)
      bool finished = true;
      if (jacobi_sweeps(c) == 0)
        for (int s = 0; s < A.nsteps; ++s)
          finished = finished && A.states_host[4 * s + 2].done != 0 && A.states_host[4 * s + 3].done != 0;
      if (finished) break;
      if (!spec) return fail(c, MHD_EHIP, "adjoint: a CG solve did not finish");
    }
  }
  if (lam0)
    for (int q = 0; q < 6; ++q)
      if (lam0[q]) {
        if (device) {
          double* dst[1] = {lam0[q]};
          rc = copy_fields(c, dst, &A.lam[q], 1);
        } else {
          rc = to_host_padded(c, A.lam[q], lam0[q], true);
        }
        if (rc) return rc;
      }
  double eb = 0.0;
  HIP_OR_FAIL(c, hipMemcpyAsync(&eb, A.eta_bar, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
  if (eta_bar) *eta_bar = eb;
  if (A.record) {
    A.br_steps = A.nsteps;
    A.br_dt = A.dt;
  }
  return MHD_OK;
}

// One step of the frozen map (mhd_adjoint_forward_frozen) on c->U, step s
// of the record, resistivity eta, Jacobi solves with cx, cy and `sweeps`.
int frozen_step(mhd_ctx* c, int s, double eta, double cx, double cy, int sweeps) {
  AdjointStore& A = *c->adj;
  const Tableau& T = c->tab;
  const double dt = A.br_dt, cg = dt * T.A_im4;
  const int n = (int)c->n;
  const int gr = grid_1d(c->n);
  LinArgs g = lin_args(c);
  const uint64_t* w1 = A.br + 2LL * s * c->n;
  const uint64_t* w2 = w1 + c->n;
  int rc;
#define KL(...) do { rc = launch(c, 8, [&] { hipLaunchKernelGGL(__VA_ARGS__); }); if (rc) return rc; } while (0)
  auto rhs = [&](double* const* st, const uint64_t* w, double* const* out) -> int {
    FrozenArgs a;
    a.nx = c->nx;
    a.ny = c->ny;
    a.idx = 1.0 / c->dx;
    a.idy = 1.0 / c->dy;
    a.br = w;
    for (int q = 0; q < 6; ++q) {
      a.s[q] = st[q];
      a.P[q] = A.P[q];
      a.out[q] = out[q];
    }
    for (int q = 0; q < 5; ++q) {
      a.fx[q] = A.xrb[q];
      a.fy[q] = A.yrb[q];
    }
    int r;
    if ((r = launch(c, 8, [&] { hipLaunchKernelGGL(kf_prims, dim3(gr), dim3(256), 0, c->stream, a); }))) return r;
    if ((r = launch(c, 8, [&] { hipLaunchKernelGGL(kf_faces, dim3(gr), dim3(256), 0, c->stream, a); }))) return r;
    return launch(c, 8, [&] { hipLaunchKernelGGL(kf_update, dim3(gr), dim3(256), 0, c->stream, a); });
  };
  // stage 1: f1 = L(U); S* = U + dt gamma f1, W = U + dt delta f1
  if ((rc = rhs(c->U, w1, A.fb))) return rc;
  KL(kf_comb1, dim3(gr), dim3(256), 0, c->stream, n, f6(c->U), f6(A.fb), dt * T.A_ex3, dt * T.A_ex6, w1, f6(A.sb),
     f6(A.wb));
  // implicit 1: x1 = A^-1 eta J(S*.B); S.B = S*.B + cg R(x1)
  KL(kf_J, dim3(gr), dim3(256), 0, c->stream, g, (const double*)A.sb[4], (const double*)A.sb[5], eta, A.q1);
  if ((rc = jacobi_run(c, cx, cy, sweeps, A.q1))) return rc;
  KL(ka_R, dim3(gr), dim3(256), 0, c->stream, g, (const double*)A.sb[4], (const double*)A.sb[5], (const double*)c->x,
     (const CgState*)nullptr, cg, A.sb[4], A.sb[5]);
  // stage 2: U3* = W + dt (1 - delta) L(S) + dt (1 - gamma) R(eta J(S.B)) on B
  if ((rc = rhs(A.sb, w2, A.fb))) return rc;
  KL(kf_J, dim3(gr), dim3(256), 0, c->stream, g, (const double*)A.sb[4], (const double*)A.sb[5], eta, A.q1);
  KL(kf_comb2, dim3(gr), dim3(256), 0, c->stream, g, f6(A.wb), f6(A.fb), (const double*)A.q1, dt * T.A_ex7,
     dt * T.A_im7, w2, f6(c->U));
  // implicit 2: B3 = U3*.B + cg R(x2), x2 = A^-1 eta J(U3*.B); U'.B = U3*.B + cg R(eta J(B3))
  KL(kf_J, dim3(gr), dim3(256), 0, c->stream, g, (const double*)c->U[4], (const double*)c->U[5], eta, A.q1);
  if ((rc = jacobi_run(c, cx, cy, sweeps, A.q1))) return rc;
  KL(ka_R, dim3(gr), dim3(256), 0, c->stream, g, (const double*)c->U[4], (const double*)c->U[5], (const double*)c->x,
     (const CgState*)nullptr, cg, A.gx, A.gy);
  KL(kf_J, dim3(gr), dim3(256), 0, c->stream, g, (const double*)A.gx, (const double*)A.gy, eta, A.q1);
  KL(ka_R, dim3(gr), dim3(256), 0, c->stream, g, (const double*)c->U[4], (const double*)c->U[5], (const double*)A.q1,
     (const CgState*)nullptr, cg, c->U[4], c->U[5]);
#undef KL
  return MHD_OK;
}

}  // namespace

extern "C" {

int mhd_adjoint_backward(mhd_ctx* c, const double* const* lam_final, double* const* lam0, double* eta_bar) {
  return backward_entry(c, lam_final, lam0, eta_bar, false);
}

int mhd_adjoint_backward_device(mhd_ctx* c, const double* const* lam_final, double* const* lam0, double* eta_bar) {
  return backward_entry(c, lam_final, lam0, eta_bar, true);
}

int mhd_explicit_rhs_vjp(mhd_ctx* c, const double* const* lam, double* const* out) {
  return rhs_vjp_entry(c, lam, out, true);
}

int mhd_adjoint_record_branches(mhd_ctx* c, int on) {
  if (check(c)) return MHD_EINVAL;
  int rc = ensure_store(c, 0, 0);
  if (rc) return rc;
  c->adj->record = (on != 0);
  return MHD_OK;
}

int mhd_adjoint_use_graph(mhd_ctx* c, int on) {
  if (check(c)) return MHD_EINVAL;
  int rc = ensure_store(c, 0, 0);
  if (rc) return rc;
  AdjointStore& A = *c->adj;
  A.use_graph = (on != 0);
  if (!A.use_graph && A.graph) {
    (void)hipGraphExecDestroy(A.graph);
    A.graph = nullptr;
  }
  return MHD_OK;
}

int mhd_adjoint_forward_frozen(mhd_ctx* c, int nsteps, double eta) {
  if (check(c)) return MHD_EINVAL;
  if (!c->adj || c->adj->br_steps < 1)
    return fail(c, MHD_EINVAL, "mhd_adjoint_forward_frozen: no branch record (mhd_adjoint_record_branches, then "
                               "mhd_adjoint_backward)");
  AdjointStore& A = *c->adj;
  if (nsteps < 1 || nsteps > A.br_steps)
    return fail(c, MHD_EINVAL, "mhd_adjoint_forward_frozen: need 1 <= nsteps <= the recorded steps");
  if (eta <= 0) eta = c->eta;
  if (!std::isfinite(eta)) return fail(c, MHD_EINVAL, "mhd_adjoint_forward_frozen: eta must be finite");
  const double alpha = (A.br_dt * eta) * c->tab.A_im4;
  const double cx = alpha / (c->dx * c->dx), cy = alpha / (c->dy * c->dy);
  const int sweeps = jacobi_sweep_count(cx, cy);
  if (sweeps < 1)
    return fail(c, MHD_EINVAL, "mhd_adjoint_forward_frozen: the implicit operator is too far from the identity for "
                               "exact Jacobi solves (2 (cx + cy) / (1 + 2 (cx + cy)) > 1/4)");
  for (int s = 0; s < nsteps; ++s) {
    int rc = frozen_step(c, s, eta, cx, cy, sweeps);
    if (rc) return rc;
  }
  if (int rc = reset_raw_density(c)) return rc;
  HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
  return MHD_OK;
}

int mhd_explicit_rhs_vjp_unfused(mhd_ctx* c, const double* const* lam, double* const* out) {
  return rhs_vjp_entry(c, lam, out, false);
}

}  // extern "C"
