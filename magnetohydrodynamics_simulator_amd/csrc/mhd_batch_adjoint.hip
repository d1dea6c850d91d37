// mhd_batch_adjoint.hip -- discrete adjoint of a batched ensemble (C ABI
// mhd_batch_adjoint_*): the adjoint of mhd_adjoint.hip for every run of a
// batch at once, each kernel launched once for all runs.
//
// Every run takes n steps of its own fixed time step (mhd_batch_set_schedule
// dt_fixed) and needs eta > 0.  The forward pass writes each step's stage data
// straight into a chain of pooled arrays (no copies): the state before the
// step Un[k], the first-stage state S[k] with S*.B in SsB[k], the second-stage
// field U3*.B in U3sB[k], the state after the step Un[k + 1]; the implicit
// solutions x1, x2 of every step and their final solver states are kept for
// all steps.  The chain holds one segment of `every` steps; a state copy is
// kept at the start of each segment.  The backward pass walks the segments
// from the last: a segment whose stage data the chain does not hold is
// recomputed from its checkpoint with the stored implicit solutions (no
// solves), then its adjoint steps run backwards.  The chain of the last
// segment still holds the forward data, so with every >= n nothing is
// recomputed.
//
// Per-cell kernels run over a table of (run, 256-cell block) jobs; the eta
// derivative's inner products leave one partial per job and a per-run kernel
// sums a run's partials in a fixed order.  The adjoint solves use, per run,
// LDS-tiled Jacobi sweeps where the operator is close enough to the identity
// (the rule of the single-run adjoint) and CG for the other runs: all CG runs
// in one batched solve, first speculatively (no host round trip; the longest
// forward solve of the step plus a margin), verified after the sweep, and
// repeated with host-checked solves if a run did not converge.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mhd_mi355x.h"
#include "mhd_adjoint_lin.hpp"
#include "mhd_adjoint_march.hpp"
#include "mhd_adjoint_math.hpp"
#include "mhd_batch.hpp"
#include "mhd_kernels.hpp"

using namespace mhd;
using namespace mhd_batch_host;

namespace mhd {

constexpr int kCellBlock = 256;   // cells per job of the per-cell kernels
constexpr int kMarchRows = 64;    // rows per segment of the fused adjoint of L

struct BArgs {
  const RunGeom* geom;
  const RunState* rs;
  const int2* jobs;   // (run, block of kCellBlock cells within the run)
};

struct BCell {
  int run, vb, o, i, j;
  bool in;
  long long base;
  LinArgs g;
};

__device__ __forceinline__ BCell bcell(const BArgs& b) {
  BCell c;
  c.vb = xcd_block(blockIdx.x, gridDim.x);
  const int2 job = b.jobs[c.vb];
  c.run = job.x;
  const RunGeom& g = b.geom[c.run];
  c.base = g.base;
  c.g.nx = g.nx;
  c.g.ny = g.ny;
  c.g.idx = g.inv_dx;
  c.g.idy = g.inv_dy;
  c.o = job.y * kCellBlock + threadIdx.x;
  c.in = c.o < g.nx * g.ny;
  c.j = c.in ? c.o / g.nx : 0;
  c.i = c.in ? c.o - c.j * g.nx : 0;
  return c;
}

__device__ __forceinline__ Fields6 shift(const Fields6& f, long long base) {
  Fields6 r;
#pragma unroll
  for (int q = 0; q < 6; ++q) r.f[q] = f.f[q] + base;
  return r;
}

// (fx, fy) = (bx, by) + dt fdt R(x); st: the per-run final states of the solve of x
__global__ __launch_bounds__(256) void kba_R(BArgs b, const double* bx, const double* by, const double* x,
                                             const CgState* st, double fdt, double* fx, double* fy) {
  const BCell c = bcell(b);
  if (!c.in) return;
  R_cell(c.g, bx + c.base, by + c.base, x + c.base, x_is_zero(st + c.run), b.rs[c.run].dt * fdt, fx + c.base,
         fy + c.base, c.i, c.j);
}

// q = dt fdt R^T(fx, fy)
__global__ __launch_bounds__(256) void kba_Rt(BArgs b, const double* fx, const double* fy, double fdt, double* q) {
  const BCell c = bcell(b);
  if (!c.in) return;
  Rt_cell(c.g, fx + c.base, fy + c.base, b.rs[c.run].dt * fdt, q + c.base, c.i, c.j);
}

// (fx, fy) += eta J^T(q)
__global__ __launch_bounds__(256) void kba_Jt(BArgs b, const double* q, double* fx, double* fy) {
  const BCell c = bcell(b);
  if (!c.in) return;
  Jt_cell(c.g, q + c.base, b.geom[c.run].eta, fx + c.base, fy + c.base, c.i, c.j);
}

// t = eta J^T(q), out = base + t
__global__ __launch_bounds__(256) void kba_Jt2(BArgs b, const double* q, const double* bx, const double* by,
                                               double* tx, double* ty, double* ox, double* oy) {
  const BCell c = bcell(b);
  if (!c.in) return;
  const long long o = c.base;
  Jt2_cell(c.g, q + o, b.geom[c.run].eta, bx + o, by + o, tx + o, ty + o, ox + o, oy + o, c.i, c.j);
}

// partial[job] = block sum of q . J(bx, by)
__global__ __launch_bounds__(256) void kba_dot_J(BArgs b, const double* q, const double* bx, const double* by,
                                                 double* partial) {
  const BCell c = bcell(b);
  const double acc = c.in ? q[c.base + c.o] * corner_J(c.g, bx + c.base, by + c.base, c.i, c.j) : 0.0;
  block_partial(acc, partial + c.vb);
}

// partial[job] = block sum of u . Lap(v) (0 where v solved with no product)
__global__ __launch_bounds__(256) void kba_dot_lap(BArgs b, const double* u, const double* v, const CgState* st,
                                                   double* partial) {
  const BCell c = bcell(b);
  const double acc =
      (c.in && !x_is_zero(st + c.run)) ? lap_dot_cell(c.g, u + c.base, v + c.base, c.i, c.j) : 0.0;
  block_partial(acc, partial + c.vb);
}

__global__ __launch_bounds__(256) void kba_stage2_in(BArgs b, Fields6 lam, Fields6 wb, Fields6 fb, Fields6 sb,
                                                     const double* rho_next, double fdt) {
  const BCell c = bcell(b);
  if (!c.in) return;
  stage2_in_cell(c.o, shift(lam, c.base), shift(wb, c.base), shift(fb, c.base), shift(sb, c.base),
                 rho_next + c.base, b.rs[c.run].dt * fdt);
}

__global__ __launch_bounds__(256) void kba_stage1_out(BArgs b, Fields6 sb, Fields6 wb, Fields6 fb, Fields6 lam,
                                                      const double* rho_s, double fs, double fw) {
  const BCell c = bcell(b);
  if (!c.in) return;
  const double dt = b.rs[c.run].dt;
  stage1_out_cell(c.o, shift(sb, c.base), shift(wb, c.base), shift(fb, c.base), shift(lam, c.base),
                  rho_s + c.base, dt * fs, dt * fw);
}

// The fused adjoint of the explicit right-hand side (mhd_adjoint_march.hpp)
// of every run: jobs (run, strip + segment x strips), pooled arrays.
struct BMarchArgs {
  const RunGeom* geom;
  const int2* jobs;
  MarchArgs p;          // pooled pointers
  int rows;             // rows per segment
};

__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(MHD_MARCH_WAVES))) void kba_march(BMarchArgs b) {
  const int2 job = b.jobs[blockIdx.x];
  const RunGeom& g = b.geom[job.x];
  MarchArgs a = b.p;
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    a.s[q] += g.base;
    a.lam[q] += g.base;
    a.out[q] += g.base;
  }
  a.nx = g.nx;
  a.ny = g.ny;
  a.rows = b.rows;
  a.idx = g.inv_dx;
  a.idy = g.inv_dy;
  const int strips = (g.nx + kAdjOut - 1) / kAdjOut;
  adj_march(a, job.y % strips, job.y / strips);
}

// Launch l of the Jacobi sweeps of every run with more than l launches
// (launches[run]; 0 = the run solves by CG).  Per run the launches ping-pong
// between jac and x so that the run's last launch writes x.
struct JacArgs {
  const RunGeom* geom;
  const RunState* rs;
  const int2* jobs;     // (run, Jacobi tile within the run), tiles x fastest
  const int* launches;
};

__global__ __launch_bounds__(256) void kba_jacobi(JacArgs b, int l, const double* rhs, double* x, double* jac) {
  const int2 job = b.jobs[blockIdx.x];
  const int L = b.launches[job.x];
  if (l >= L) return;
  const RunGeom& g = b.geom[job.x];
  const RunState& s = b.rs[job.x];
  const LinArgs lg{g.nx, g.ny, g.inv_dx, g.inv_dy};
  const bool to_x = ((L - 1 - l) & 1) == 0;
  double* out = (to_x ? x : jac) + g.base;
  const double* in = (l == 0) ? nullptr : (to_x ? jac : x) + g.base;
  const int tiles_x = (g.nx + kJacTile - 1) / kJacTile;
  jacobi_tile<kJacTile, kJacSweeps>(lg, s.cx, s.cy, s.invd, rhs + g.base, in, out, job.y % tiles_x,
                                    job.y / tiles_x);
}

// Fresh solver states: CG runs start, Jacobi runs are marked done.
__global__ void kba_cg_reset(const RunGeom* geom, const int* launches, CgState* st, int nruns) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nruns) return;
  CgState c{};
  c.max_iters = geom[r].max_iters;
  if (launches[r] != 0) c.done = 3;
  st[r] = c;
}

// x = 0 for CG runs whose solve did no product
__global__ __launch_bounds__(256) void kba_zero_if_no_product(BArgs b, const int* launches, const CgState* st,
                                                              double* x) {
  const BCell c = bcell(b);
  if (!c.in || launches[c.run] != 0 || st[c.run].spmv != 0) return;
  x[c.base + c.o] = 0.0;
}

// eta_bar[run] += sum_k scale_k * (sum of the run's partials of slot k), in a
// fixed order; scales (1, 1, dt gamma, 1, 1, dt gamma).  One block per run.
__global__ __launch_bounds__(256) void kba_accumulate(const double* partial, int ncj, const int2* range,
                                                      const RunState* rs, double a_im4, double* eta_bar) {
  __shared__ double lds[8];
  const int run = blockIdx.x;
  const int2 rg = range[run];
  const double dg = rs[run].dt * a_im4;
  double e = 0.0;
  for (int k = 0; k < 6; ++k) {
    double acc = 0.0;
    for (int j = rg.x + threadIdx.x; j < rg.y; j += blockDim.x) acc += partial[(long long)k * ncj + j];
    e += ((k == 2 || k == 5) ? dg : 1.0) * block_total(acc, lds);
  }
  if (threadIdx.x == 0) eta_bar[run] += e;
}

__global__ __launch_bounds__(256) void kba_fill(long long n, double* y, double v) {
  for (long long o = blockIdx.x * (long long)blockDim.x + threadIdx.x; o < n; o += (long long)gridDim.x * blockDim.x)
    y[o] = v;
}

}  // namespace mhd

// ---------------------------------------------------------------------------
// Host side.
// ---------------------------------------------------------------------------

namespace mhd_batch_host {

using F6 = std::array<double*, 6>;
using F2 = std::array<double*, 2>;

struct BatchAdjoint {
  int nsteps = 0, every = 0, seg = 0, nseg = 0;
  int chain_seg = -1;               // segment whose stage data the chain holds
  std::vector<double*> owned;       // every pooled device array of this store
  std::vector<F6> Un;               // seg + 1 states of the chain
  std::vector<F6> S;                // seg first-stage states (S.B after the solve)
  std::vector<F2> SsB, U3sB;        // seg
  std::vector<F6> ckpt;             // state at the start of each segment
  std::vector<double*> xs;          // the two implicit solutions of every step
  CgState* fst = nullptr;           // forward solve states [2 nsteps][nruns]
  CgState* ast = nullptr;           // adjoint solve states [2 nsteps][nruns]
  CgState* ast_host = nullptr;      // pinned mirror
  RunState* rs = nullptr;           // driver state of the recorded steps
  std::vector<int> fwd_spmv;        // longest forward solve over the runs, 2 per step
  std::vector<int> jl_h;            // Jacobi launches per run (0: CG)
  int* jl = nullptr;
  int max_jl = 0;
  bool any_cg = false;
  int2* cjobs = nullptr;            // per-cell jobs
  int ncj = 0;
  int2* crange = nullptr;           // [first, end) of each run's jobs
  int2* jjobs = nullptr;            // Jacobi tiles of the Jacobi runs
  int njj = 0;
  int2* mjobs = nullptr;            // (run, strip + segment x strips) of the fused adjoint of L
  int nmj = 0;
  double* lam[6] = {};              // adjoint of the state after the step (in), before it (out)
  double* lam_in[6] = {};           // the final-state adjoint, kept for a repeated sweep
  double* wb[6] = {};
  double* sb[6] = {};
  double* fb[6] = {};
  double* q1 = nullptr;
  double* fx = nullptr;
  double* fy = nullptr;
  double* gx = nullptr;
  double* gy = nullptr;
  double* jac = nullptr;
  double* partial = nullptr;        // 6 slots of ncj job partials
  double* eta_bar = nullptr;        // one per run
  std::vector<double> eta_bar_h;
  bool recorded = false;            // a forward pass completed
  bool have_result = false;
};

void destroy_batch_adjoint(mhd_batch* b) {
  if (!b || !b->adj) return;
  BatchAdjoint* A = b->adj;
  for (double* p : A->owned)
    if (p) (void)hipFree(p);
  void* bufs[] = {A->fst, A->ast, A->rs, A->jl, A->cjobs, A->crange, A->jjobs, A->mjobs, A->partial, A->eta_bar};
  for (void* p : bufs)
    if (p) (void)hipFree(p);
  if (A->ast_host) (void)hipHostFree(A->ast_host);
  delete A;
  b->adj = nullptr;
}

}  // namespace mhd_batch_host

namespace {

int grid_of(long long n) {
  long long g = (n + 255) / 256;
  return (int)std::max(1LL, std::min(g, 8192LL));
}

int field(mhd_batch* b, double** p) {
  if (hipMalloc(p, (size_t)b->total * sizeof(double)) != hipSuccess) {
    *p = nullptr;
    return bfail(b, MHD_ENOMEM, "batch adjoint: device allocation failed");
  }
  b->adj->owned.push_back(*p);
  return MHD_OK;
}

template <typename T>
int dev_alloc(mhd_batch* b, T** p, size_t count) {
  if (hipMalloc(p, sizeof(T) * std::max<size_t>(count, 1)) != hipSuccess) {
    *p = nullptr;
    return bfail(b, MHD_ENOMEM, "batch adjoint: device allocation failed");
  }
  return MHD_OK;
}

int copy_pooled(mhd_batch* b, double* const* dst, double* const* src, int nf) {
  CopyArgs a{};
  for (int q = 0; q < nf; ++q) {
    a.src[q] = src[q];
    a.dst[q] = dst[q];
  }
  a.nf = nf;
  BLAUNCH(b, k_copy_fields, dim3(grid_of(b->total)), dim3(256), 0, b->stream, a, (long long)b->total);
  return MHD_OK;
}

int fill_pooled(mhd_batch* b, double* y, long long n, double v) {
  BLAUNCH(b, kba_fill, dim3(grid_of(n)), dim3(256), 0, b->stream, n, y, v);
  return MHD_OK;
}

// Jacobi launches of a run's adjoint solves (0: CG): the sweep count of
// jacobi_sweep_count for the operator kb_dt builds for the run's fixed step
// (alpha = dt eta gamma), 8 sweeps per launch.
int jacobi_launches(const RunGeom& g, double a_im4) {
  const double alpha = (g.dt_fixed * g.eta) * a_im4;
  const int sweeps = jacobi_sweep_count(alpha / (g.dx * g.dx), alpha / (g.dy * g.dy));
  return (sweeps + kJacSweeps - 1) / kJacSweeps;
}

// The store for n steps with checkpoints every k steps.
int build_store(mhd_batch* b, int nsteps, int every) {
  b->adj = new BatchAdjoint();
  BatchAdjoint& A = *b->adj;
  A.nsteps = nsteps;
  A.every = every;
  A.seg = std::min(every, nsteps);
  A.nseg = (nsteps + every - 1) / every;
  int rc;
  auto f6 = [&](F6& f) -> int {
    for (double*& p : f)
      if (int r = field(b, &p)) return r;
    return MHD_OK;
  };
  A.Un.resize(A.seg + 1);
  A.S.resize(A.seg);
  A.SsB.resize(A.seg);
  A.U3sB.resize(A.seg);
  A.ckpt.resize(A.nseg);
  for (F6& f : A.Un)
    if ((rc = f6(f))) return rc;
  for (int k = 0; k < A.seg; ++k) {
    if ((rc = f6(A.S[k]))) return rc;
    for (int q = 0; q < 2; ++q) {
      if ((rc = field(b, &A.SsB[k][q]))) return rc;
      if ((rc = field(b, &A.U3sB[k][q]))) return rc;
    }
  }
  for (F6& f : A.ckpt)
    if ((rc = f6(f))) return rc;
  A.xs.resize(2 * nsteps);
  for (double*& p : A.xs)
    if ((rc = field(b, &p))) return rc;
  double** groups[] = {A.lam, A.lam_in, A.wb, A.sb, A.fb};
  for (double** gp : groups)
    for (int q = 0; q < 6; ++q)
      if ((rc = field(b, &gp[q]))) return rc;
  double** singles[] = {&A.q1, &A.fx, &A.fy, &A.gx, &A.gy, &A.jac};
  for (double** p : singles)
    if ((rc = field(b, p))) return rc;
  const int nr = b->nruns;
  if ((rc = dev_alloc(b, &A.fst, (size_t)2 * nsteps * nr)) || (rc = dev_alloc(b, &A.ast, (size_t)2 * nsteps * nr)) ||
      (rc = dev_alloc(b, &A.rs, nr)) || (rc = dev_alloc(b, &A.jl, nr)) || (rc = dev_alloc(b, &A.crange, nr)) ||
      (rc = dev_alloc(b, &A.eta_bar, nr)))
    return rc;
  if (hipHostMalloc(&A.ast_host, sizeof(CgState) * 2 * nsteps * nr) != hipSuccess) {
    A.ast_host = nullptr;
    return bfail(b, MHD_ENOMEM, "batch adjoint: pinned allocation failed");
  }
  // per-cell jobs and the per-run ranges
  std::vector<int2> cj, cr(nr);
  for (int r = 0; r < nr; ++r) {
    const RunGeom& g = b->geom_h[r];
    const long long n = (long long)g.nx * g.ny;
    cr[r].x = (int)cj.size();
    for (long long k = 0; k < (n + kCellBlock - 1) / kCellBlock; ++k) cj.push_back(make_int2(r, (int)k));
    cr[r].y = (int)cj.size();
  }
  A.ncj = (int)cj.size();
  std::vector<int2> mj;
  for (int r = 0; r < nr; ++r) {
    const RunGeom& g = b->geom_h[r];
    const int w = ((g.nx + kAdjOut - 1) / kAdjOut) * ((g.ny + kMarchRows - 1) / kMarchRows);
    for (int k = 0; k < w; ++k) mj.push_back(make_int2(r, k));
  }
  A.nmj = (int)mj.size();
  if ((rc = dev_alloc(b, &A.cjobs, cj.size())) || (rc = dev_alloc(b, &A.partial, (size_t)6 * A.ncj)) ||
      (rc = dev_alloc(b, &A.mjobs, mj.size())))
    return rc;
  BHIP(b, hipMemcpyAsync(A.mjobs, mj.data(), sizeof(int2) * mj.size(), hipMemcpyHostToDevice, b->stream));
  BHIP(b, hipMemcpyAsync(A.cjobs, cj.data(), sizeof(int2) * cj.size(), hipMemcpyHostToDevice, b->stream));
  BHIP(b, hipMemcpyAsync(A.crange, cr.data(), sizeof(int2) * nr, hipMemcpyHostToDevice, b->stream));
  BHIP(b, hipStreamSynchronize(b->stream));
  return MHD_OK;
}

// The store is kept while n and k stay; a failed allocation leaves none.
int ensure_store(mhd_batch* b, int nsteps, int every) {
  if (b->adj && b->adj->nsteps == nsteps && b->adj->every == every) return MHD_OK;
  destroy_batch_adjoint(b);
  const int rc = build_store(b, nsteps, every);
  if (rc) destroy_batch_adjoint(b);
  return rc;
}

// The per-run solver choice of the adjoint solves and the Jacobi job table
// (the fixed steps may change between forward passes).
int plan_solves(mhd_batch* b) {
  BatchAdjoint& A = *b->adj;
  const int nr = b->nruns;
  A.jl_h.assign(nr, 0);
  A.max_jl = 0;
  A.any_cg = false;
  std::vector<int2> jj;
  for (int r = 0; r < nr; ++r) {
    const RunGeom& g = b->geom_h[r];
    A.jl_h[r] = jacobi_launches(g, b->tab.A_im4);
    A.max_jl = std::max(A.max_jl, A.jl_h[r]);
    if (A.jl_h[r] == 0) {
      A.any_cg = true;
      continue;
    }
    const int tiles = ((g.nx + kJacTile - 1) / kJacTile) * ((g.ny + kJacTile - 1) / kJacTile);
    for (int t = 0; t < tiles; ++t) jj.push_back(make_int2(r, t));
  }
  if ((int)jj.size() > A.njj) {
    if (A.jjobs) (void)hipFree(A.jjobs);
    A.jjobs = nullptr;
    A.njj = 0;
    if (int rc = dev_alloc(b, &A.jjobs, jj.size())) return rc;
  }
  A.njj = (int)jj.size();
  if (!jj.empty())
    BHIP(b, hipMemcpyAsync(A.jjobs, jj.data(), sizeof(int2) * jj.size(), hipMemcpyHostToDevice, b->stream));
  BHIP(b, hipMemcpyAsync(A.jl, A.jl_h.data(), sizeof(int) * nr, hipMemcpyHostToDevice, b->stream));
  BHIP(b, hipStreamSynchronize(b->stream));
  return MHD_OK;
}

// Step s of the recorded sequence at chain position k: the batched step of
// mhd_batch_step with its stage data written into the chain.  given: the
// recompute of the backward pass (no driver update, the stored implicit
// solutions instead of the solves).
int record_step(mhd_batch* b, int s, int k, bool given) {
  BatchAdjoint& A = *b->adj;
  const BatchTableau& T = b->tab;
  const int nr = b->nruns, nt = b->n_tile_jobs;
  const F6& in = A.Un[k];
  const F6& out = A.Un[k + 1];
  CgState* st0 = A.fst + (size_t)(2 * s) * nr;
  CgState* st1 = st0 + nr;
  double* x1 = A.xs[2 * s];
  double* x2 = A.xs[2 * s + 1];
  const RunState* rs = given ? A.rs : b->rs;
  if (!given) {
    BatchDiagArgs d{};
    d.geom = b->geom;
    d.rs = b->rs;
    d.jobs = b->tile_jobs;
    for (int q = 0; q < 6; ++q) d.s[q] = in[q];
    d.part = b->part;
    BLAUNCH(b, kb_cfl, dim3(nt), dim3(kTileX), 0, b->stream, d);
    BLAUNCH(b, kb_dt, dim3(nr), dim3(256), 0, b->stream, (const RunGeom*)b->geom, b->rs, (const double*)b->part, st0,
            st1, nr, step_factors(T), 0, nullptr);
    BHIP(b, hipMemcpyAsync(b->rs_host, b->rs, sizeof(RunState) * nr, hipMemcpyDeviceToHost, b->stream));
  }
  // stage 1: W = U + dt delta f1, S* = U + dt gamma f1 (S*.B into SsB)
  BatchStageArgs s1{};
  s1.geom = b->geom;
  s1.rs = rs;
  s1.jobs = b->stage_jobs;
  for (int q = 0; q < 6; ++q) {
    s1.s[q] = in[q];
    s1.o1[q] = b->W[q];
    s1.o2[q] = (q < 4) ? A.S[k][q] : A.SsB[k][q - 4];
  }
  BLAUNCH(b, kb_stage<1>, dim3(b->n_stage_jobs), dim3(64), 0, b->stream, s1);
  if (!given) {
    CgCall cc;
    cc.slot = 0;
    cc.bx = A.SsB[k][0];
    cc.by = A.SsB[k][1];
    cc.x = x1;
    cc.st = st0;
    cc.rs = b->rs;
    int most = 0;
    if (int rc = cg_batch(b, cc, &most)) return rc;
    A.fwd_spmv[2 * s] = most;
  }
  BLAUNCH(b, kb_finalize1, dim3(nt), dim3(kTileX), 0, b->stream,
          fin_args(b, rs, x1, st0, A.SsB[k][0], A.SsB[k][1], A.S[k][4], A.S[k][5]));
  // stage 2: U3* = W + dt (1 - delta) L(S) + dt (1 - gamma) g2 (U3*.B into U3sB)
  BatchStageArgs s2{};
  s2.geom = b->geom;
  s2.rs = rs;
  s2.jobs = b->stage_jobs;
  for (int q = 0; q < 6; ++q) {
    s2.s[q] = A.S[k][q];
    s2.u[q] = b->W[q];
    s2.o1[q] = (q < 4) ? out[q] : A.U3sB[k][q - 4];
  }
  BLAUNCH(b, kb_stage<2>, dim3(b->n_stage_jobs), dim3(64), 0, b->stream, s2);
  if (!given) {
    CgCall cc;
    cc.slot = 1;
    cc.bx = A.U3sB[k][0];
    cc.by = A.U3sB[k][1];
    cc.x = x2;
    cc.st = st1;
    cc.rs = b->rs;
    int most = 0;
    if (int rc = cg_batch(b, cc, &most)) return rc;
    A.fwd_spmv[2 * s + 1] = most;
  }
  BLAUNCH(b, kb_finalize2<false>, dim3(nt), dim3(kTileX), 0, b->stream,
          fin_args(b, rs, x2, st1, A.U3sB[k][0], A.U3sB[k][1], out[4], out[5]));
  if (!given) {
    BHIP(b, hipStreamSynchronize(b->stream));
    for (int r = 0; r < nr; ++r) {
      if (b->rs_host[r].nonfinite)
        return bfail(b, MHD_ENONFINITE, "run " + std::to_string(r) + ": non-finite wave speed");
      if (!b->rs_host[r].active)
        return bfail(b, MHD_EINVAL, "run " + std::to_string(r) + ": reached its t_end or step limit during the "
                                    "recorded steps");
    }
  }
  return MHD_OK;
}

BArgs bargs(mhd_batch* b) {
  BatchAdjoint& A = *b->adj;
  return BArgs{b->geom, A.rs, A.cjobs};
}

Fields6 six(double* const* p) {
  Fields6 f;
  for (int q = 0; q < 6; ++q) f.f[q] = p[q];
  return f;
}

// out += L'(s)^T lam for the runs' states s.
int rhs_vjp(mhd_batch* b, double* const* s, double* const* lam, double* const* out) {
  BatchAdjoint& A = *b->adj;
  BMarchArgs m{};
  m.geom = b->geom;
  m.jobs = A.mjobs;
  for (int q = 0; q < 6; ++q) {
    m.p.s[q] = s[q];
    m.p.lam[q] = lam[q];
    m.p.out[q] = out[q];
  }
  m.rows = kMarchRows;
  BLAUNCH(b, kba_march, dim3(A.nmj), dim3(64), 0, b->stream, m);
  return MHD_OK;
}

// b->x = A^-1 rhs per run (Jacobi or CG); states of the CG runs to st.
int adjoint_solve(mhd_batch* b, const double* rhs, CgState* st, int spec) {
  BatchAdjoint& A = *b->adj;
  const int nr = b->nruns;
  if (A.max_jl > 0) {
    const JacArgs ja{b->geom, A.rs, A.jjobs, A.jl};
    for (int l = 0; l < A.max_jl; ++l)
      BLAUNCH(b, kba_jacobi, dim3(A.njj), dim3(256), 0, b->stream, ja, l, rhs, b->x, A.jac);
  }
  BLAUNCH(b, kba_cg_reset, dim3((nr + 255) / 256), dim3(256), 0, b->stream, (const RunGeom*)b->geom,
          (const int*)A.jl, st, nr);
  if (!A.any_cg) return MHD_OK;
  CgCall cc;
  cc.slot = 2;
  cc.bin = rhs;
  cc.x = b->x;
  cc.st = st;
  cc.rs = A.rs;
  cc.spec = spec;
  if (int rc = cg_batch(b, cc)) return rc;
  BLAUNCH(b, kba_zero_if_no_product, dim3(A.ncj), dim3(256), 0, b->stream, bargs(b), (const int*)A.jl,
          (const CgState*)st, b->x);
  return MHD_OK;
}

// The adjoint of step s (chain position k): lam (after the step) -> lam
// (before it), the eta derivative's terms into eta_bar.  The sequence of the
// single-run adjoint_step (mhd_adjoint.hip) over all runs.
int adjoint_step(mhd_batch* b, int k, int s, int spec) {
  BatchAdjoint& A = *b->adj;
  const BatchTableau& T = b->tab;
  const int nr = b->nruns;
  const BArgs ba = bargs(b);
  const dim3 G(A.ncj), B(256);
  const long long nj = A.ncj;
  const CgState* fs0 = A.fst + (size_t)(2 * s) * nr;
  const CgState* fs1 = fs0 + nr;
  const double* x1 = A.xs[2 * s];
  const double* x2 = A.xs[2 * s + 1];
  double* P = A.partial;
  double** lam = A.lam;
  double* xa = b->x;
  // final recombination: U' = (U3*.h, U3*.B + cg R(eta J(B3))), B3 = U3*.B + cg R(x2)
  BLAUNCH(b, kba_R, G, B, 0, b->stream, ba, (const double*)A.U3sB[k][0], (const double*)A.U3sB[k][1], x2, fs1,
          T.A_im4, A.fx, A.fy);
  BLAUNCH(b, kba_Rt, G, B, 0, b->stream, ba, (const double*)lam[4], (const double*)lam[5], T.A_im4, A.q1);
  BLAUNCH(b, kba_Jt2, G, B, 0, b->stream, ba, (const double*)A.q1, (const double*)lam[4], (const double*)lam[5], A.gx,
          A.gy, A.wb[4], A.wb[5]);
  BLAUNCH(b, kba_dot_J, G, B, 0, b->stream, ba, (const double*)A.q1, (const double*)A.fx, (const double*)A.fy, P);
  BLAUNCH(b, kba_Rt, G, B, 0, b->stream, ba, (const double*)A.gx, (const double*)A.gy, T.A_im4, A.q1);
  if (int rc = adjoint_solve(b, A.q1, A.ast + (size_t)(2 * s + 1) * nr, spec)) return rc;
  BLAUNCH(b, kba_Jt, G, B, 0, b->stream, ba, (const double*)xa, A.wb[4], A.wb[5]);
  BLAUNCH(b, kba_dot_J, G, B, 0, b->stream, ba, (const double*)xa, (const double*)A.U3sB[k][0],
          (const double*)A.U3sB[k][1], P + 1 * nj);
  BLAUNCH(b, kba_dot_lap, G, B, 0, b->stream, ba, (const double*)xa, x2, fs1, P + 2 * nj);
  // stage 2
  BLAUNCH(b, kba_stage2_in, G, B, 0, b->stream, ba, six(lam), six(A.wb), six(A.fb), six(A.sb),
          (const double*)A.Un[k + 1][0], T.A_ex7);
  BLAUNCH(b, kba_Rt, G, B, 0, b->stream, ba, (const double*)A.wb[4], (const double*)A.wb[5], T.A_im7, A.q1);
  BLAUNCH(b, kba_Jt, G, B, 0, b->stream, ba, (const double*)A.q1, A.sb[4], A.sb[5]);
  BLAUNCH(b, kba_dot_J, G, B, 0, b->stream, ba, (const double*)A.q1, (const double*)A.S[k][4],
          (const double*)A.S[k][5], P + 3 * nj);
  if (int rc = rhs_vjp(b, A.S[k].data(), A.fb, A.sb)) return rc;
  // implicit 1: S.B = S*.B + cg R(x1), x1 = A^-1 eta J(S*.B)
  BLAUNCH(b, kba_Rt, G, B, 0, b->stream, ba, (const double*)A.sb[4], (const double*)A.sb[5], T.A_im4, A.q1);
  if (int rc = adjoint_solve(b, A.q1, A.ast + (size_t)(2 * s) * nr, spec)) return rc;
  BLAUNCH(b, kba_Jt, G, B, 0, b->stream, ba, (const double*)xa, A.sb[4], A.sb[5]);
  BLAUNCH(b, kba_dot_J, G, B, 0, b->stream, ba, (const double*)xa, (const double*)A.SsB[k][0],
          (const double*)A.SsB[k][1], P + 4 * nj);
  BLAUNCH(b, kba_dot_lap, G, B, 0, b->stream, ba, (const double*)xa, x1, fs0, P + 5 * nj);
  // stage 1
  BLAUNCH(b, kba_stage1_out, G, B, 0, b->stream, ba, six(A.sb), six(A.wb), six(A.fb), six(lam),
          (const double*)A.S[k][0], T.A_ex3, T.A_ex6);
  if (int rc = rhs_vjp(b, A.Un[k].data(), A.fb, lam)) return rc;
  BLAUNCH(b, kba_accumulate, dim3(nr), dim3(256), 0, b->stream, (const double*)P, A.ncj, (const int2*)A.crange,
          (const RunState*)A.rs, T.A_im4, A.eta_bar);
  return MHD_OK;
}

// The backward pass from lam (the final-state adjoint) to lam (the initial one).
int sweep(mhd_batch* b, bool spec) {
  BatchAdjoint& A = *b->adj;
  if (int rc = fill_pooled(b, A.eta_bar, b->nruns, 0.0)) return rc;
  for (int sg = A.nseg - 1; sg >= 0; --sg) {
    const int s0 = sg * A.every, s1 = std::min(A.nsteps, s0 + A.every);
    if (A.chain_seg != sg) {
      A.chain_seg = -1;
      if (int rc = copy_pooled(b, A.Un[0].data(), A.ckpt[sg].data(), 6)) return rc;
      for (int s = s0; s < s1; ++s)
        if (int rc = record_step(b, s, s - s0, true)) return rc;
      A.chain_seg = sg;
    }
    for (int s = s1 - 1; s >= s0; --s) {
      // the adjoint solves use the operator of the forward ones; their
      // right-hand sides differ, so a margin is added to the forward counts
      const int sp = std::max(A.fwd_spmv[2 * s], A.fwd_spmv[2 * s + 1]);
      if (int rc = adjoint_step(b, s - s0, s, spec ? sp + 2 + sp / 4 : 0)) return rc;
    }
  }
  return MHD_OK;
}

int check_adj(mhd_batch* b) {
  if (!b) return MHD_EINVAL;
  if (!b->adj || !b->adj->recorded) return bfail(b, MHD_EINVAL, "batch adjoint: no recorded forward pass");
  return MHD_OK;
}

}  // namespace

extern "C" {

int mhd_batch_adjoint_forward(mhd_batch* b, int nsteps, int checkpoint_every) {
  if (!b) return MHD_EINVAL;
  if (batch_open(b)) return bfail(b, MHD_EINVAL, "mhd_batch_adjoint_forward: the batched adjoint supports periodic runs only");
  b->cfl_fresh = false;
  if (nsteps < 1 || checkpoint_every < 1)
    return bfail(b, MHD_EINVAL, "mhd_batch_adjoint_forward: need nsteps >= 1 and checkpoint_every >= 1");
  for (int r = 0; r < b->nruns; ++r) {
    const RunGeom& g = b->geom_h[r];
    if (!(g.eta > 0) || !(g.dt_fixed > 0))
      return bfail(b, MHD_EINVAL, "run " + std::to_string(r) +
                                      ": the adjoint needs eta > 0 and a fixed time step (mhd_batch_set_schedule)");
  }
  int rc = ensure_store(b, nsteps, checkpoint_every);
  if (rc) return rc;
  BatchAdjoint& A = *b->adj;
  if ((rc = plan_solves(b))) return rc;
  A.chain_seg = -1;
  A.recorded = false;
  A.have_result = false;
  A.fwd_spmv.assign(2 * nsteps, 1);
  for (int q = 0; q < 6; ++q)   // a run whose final adjoint is not set gets zero
    if ((rc = fill_pooled(b, A.lam[q], b->total, 0.0))) return rc;
  if ((rc = copy_pooled(b, A.Un[0].data(), b->U, 6))) return rc;
  int prev_len = 0;
  for (int sg = 0; sg < A.nseg; ++sg) {
    const int s0 = sg * checkpoint_every, s1 = std::min(nsteps, s0 + checkpoint_every);
    if (sg > 0) std::swap(A.Un[0], A.Un[prev_len]);   // the last state of the chain starts the next segment
    if ((rc = copy_pooled(b, A.ckpt[sg].data(), A.Un[0].data(), 6))) return rc;
    for (int s = s0; s < s1; ++s)
      if ((rc = record_step(b, s, s - s0, false))) return rc;
    prev_len = s1 - s0;
  }
  A.chain_seg = A.nseg - 1;
  if ((rc = copy_pooled(b, b->U, A.Un[prev_len].data(), 6))) return rc;
  BHIP(b, hipMemcpyAsync(A.rs, b->rs, sizeof(RunState) * b->nruns, hipMemcpyDeviceToDevice, b->stream));
  BHIP(b, hipStreamSynchronize(b->stream));
  b->steps_taken += nsteps;
  A.recorded = true;
  return MHD_OK;
}

int mhd_batch_adjoint_set_lambda(mhd_batch* b, int run, const double* const* lam_final) {
  if (int rc = check_adj(b)) return rc;
  if (run < 0 || run >= b->nruns) return bfail(b, MHD_EINVAL, "run index out of range");
  if (!lam_final) return bfail(b, MHD_EINVAL, "lam_final is NULL");
  for (int q = 0; q < 6; ++q)
    if (!lam_final[q]) return bfail(b, MHD_EINVAL, "lam_final has a NULL field");
  for (int q = 0; q < 6; ++q)
    if (int rc = stage_in(b, run, lam_final[q], b->adj->lam[q])) return rc;
  BHIP(b, hipStreamSynchronize(b->stream));
  b->adj->have_result = false;
  return MHD_OK;
}

int mhd_batch_adjoint_backward(mhd_batch* b) {
  if (int rc = check_adj(b)) return rc;
  b->cfl_fresh = false;
  BatchAdjoint& A = *b->adj;
  const int nr = b->nruns;
  int rc = copy_pooled(b, A.lam_in, A.lam, 6);
  if (rc) return rc;
  // first pass: every adjoint CG solve speculative, verified afterwards; a
  // solve that did not finish sends the sweep round again with host-checked
  // solves
  for (int attempt = 0; attempt < 2; ++attempt) {
    const bool spec = (attempt == 0);
    if (!spec && (rc = copy_pooled(b, A.lam, A.lam_in, 6))) return rc;
    if ((rc = sweep(b, spec))) return rc;
    BHIP(b, hipMemcpyAsync(A.ast_host, A.ast, sizeof(CgState) * 2 * A.nsteps * nr, hipMemcpyDeviceToHost, b->stream));
    BHIP(b, hipStreamSynchronize(b->stream));
    bool finished = true;
    for (int k = 0; k < 2 * A.nsteps * nr && finished; ++k) finished = A.ast_host[k].done != 0;
    if (finished) break;
    if (!spec) return bfail(b, MHD_EHIP, "batch adjoint: a CG solve did not finish");
  }
  A.eta_bar_h.assign(nr, 0.0);
  BHIP(b, hipMemcpyAsync(A.eta_bar_h.data(), A.eta_bar, sizeof(double) * nr, hipMemcpyDeviceToHost, b->stream));
  BHIP(b, hipStreamSynchronize(b->stream));
  A.have_result = true;
  return MHD_OK;
}

int mhd_batch_adjoint_result(mhd_batch* b, int run, double* const* lam0, double* eta_bar) {
  if (int rc = check_adj(b)) return rc;
  if (run < 0 || run >= b->nruns) return bfail(b, MHD_EINVAL, "run index out of range");
  BatchAdjoint& A = *b->adj;
  if (!A.have_result) return bfail(b, MHD_EINVAL, "mhd_batch_adjoint_result before mhd_batch_adjoint_backward");
  if (lam0)
    for (int q = 0; q < 6; ++q)
      if (lam0[q])
        if (int rc = stage_out(b, run, A.lam[q], lam0[q])) return rc;
  if (eta_bar) *eta_bar = A.eta_bar_h[run];
  return MHD_OK;
}

}  // extern "C"
