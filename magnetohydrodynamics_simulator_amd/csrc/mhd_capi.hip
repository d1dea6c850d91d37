// mhd_capi.hip -- the C ABI (include/mhd_mi355x.h) over the gfx950 kernels.
//
// Host-side orchestration of one ARS(2,2,2) step (include/solvers.hpp:491-613
// of the reference) on device-resident compact fields:
//   stage 1       f1 = L(U); W = U + dt d f1 (held in F1), S = U + dt g f1   (1 kernel)
//   implicit 1    CG on eta J(S.B), S.B += dt g rate(x)           (start, 2/iter, 1)
//   stage 2       U <- W + dt (1-d) L(S) [+ dt (1-g) g2(S)]        (1 kernel)
//   implicit 2    CG on eta J(U.B), B^{n+1} = U.B + dt g g3(B3)   (start, 2/iter, 1)
// W is the reference's grid_whole_step after its f1 term (solvers.hpp:548-553):
// the same additions in the same order, formed in stage 1 so that stage 2
// reads one field per cell instead of U^n and f1.
// with g = (2 - sqrt 2)/2 and d = 1 - 1/(2g) (solvers.hpp:41-42).  The stage-1
// explicit resistive term g1 has weight 0 in every combination
// (A_im[3] = A_im[6] = b_im[0] = 0) and is not evaluated.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "../../include/mhd_mi355x.h"
#include "mhd_context.hpp"
#include "mhd_kernels.hpp"

using namespace mhd;
using namespace mhd_host;

namespace {
thread_local std::string g_create_error;
}  // namespace

namespace mhd_host {

int fail(mhd_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  else g_create_error = msg;
  return code;
}

int flush_profile(mhd_ctx* c) {
  HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
  for (int k = 0; k < kNumKernelIds; ++k) {
    for (auto& pr : c->pending[k]) {
      float ms = 0.f;
      HIP_OR_FAIL(c, hipEventElapsedTime(&ms, pr.first, pr.second));
      c->prof_ms[k] += ms;
      c->prof_n[k] += 1;
      (void)hipEventDestroy(pr.first);
      (void)hipEventDestroy(pr.second);
    }
    c->pending[k].clear();
  }
  return MHD_OK;
}

int grid_1d(long long n) {
  long long b = (n + 255) / 256;
  if (b > 8192) b = 8192;
  if (b < 1) b = 1;
  return (int)b;
}

StageArgs stage_args(mhd_ctx* c) {
  StageArgs a{};
  a.nx = c->nx;
  a.ny = c->ny;
  a.rows = c->stage_rows;
  a.open_x = c->o.bc_x == MHD_BC_OUTFLOW;
  a.open_y = c->o.bc_y == MHD_BC_OUTFLOW;
  a.dx = c->dx;
  a.dy = c->dy;
  a.inv_dx = 1.0 / c->dx;
  a.inv_dy = 1.0 / c->dy;
  a.eta = c->eta;
  a.rho_raw = c->rho_raw;
  a.strips = (c->nx + kStripOut - 1) / kStripOut;
  a.segs = (c->ny + a.rows - 1) / a.rows;
  a.band = kStageBands > 0 ? (a.strips + kStageBands - 1) / kStageBands : 0;
  return a;
}

dim3 stage_grid(const StageArgs& a) {
  return a.band > 0 ? dim3((unsigned)(a.strips * a.segs)) : dim3((unsigned)a.strips, (unsigned)a.segs);
}

int run_stage(mhd_ctx* c, int mode, const StageArgs& a, int id) {
  const dim3 grid = stage_grid(a);
  if (mode == 0) return launch_k(c, id, k_stage<0>, grid, dim3(64), a);
  if (mode == 1) return launch_k(c, id, k_stage<1>, grid, dim3(64), a);
  return launch_k(c, id, k_stage<2>, grid, dim3(64), a);
}

CgArgs cg_args(mhd_ctx* c, const double* bx, const double* by) {
  CgArgs a{};
  a.nx = c->nx;
  a.ny = c->ny;
  a.open_x = c->o.bc_x == MHD_BC_OUTFLOW;
  a.open_y = c->o.bc_y == MHD_BC_OUTFLOW;
  a.dx = c->dx;
  a.dy = c->dy;
  a.inv_dx = 1.0 / c->dx;
  a.inv_dy = 1.0 / c->dy;
  a.eta = c->eta;
  const double cx = c->alpha / (c->dx * c->dx);
  const double cy = c->alpha / (c->dy * c->dy);
  a.cx = cx;
  a.cy = cy;
  a.diag = 1 + 2 * (cx + cy);
  a.invd = 1.0 / a.diag;
  a.tol = c->o.cg_tol;
  a.bx = bx;
  a.by = by;
  a.r = c->r;
  a.p0 = c->p[0];
  a.x = c->x;
  a.part_s = c->part_s;
  a.part_a = c->part_a;
  a.part_b = c->part_b;
  a.st = c->st;
  a.ty = c->tile_rows;
  a.tiles_y = c->tiles_y;
  return a;
}

// One conjugate-gradient solve; the right-hand side is eta J(bx, by) or bin.
// Kernel order: start (b, r, p, first product), B(0), then A(it), B(it)
// pairs.  Each A takes the convergence decision on the partials of the B
// before it; kernels launched after the decision return at once.  The host
// reads the state back after each chunk; the first chunk is as long as the
// previous solve's product count, so a steady run syncs once per solve.
int cg_run(mhd_ctx* c, const double* bx, const double* by, const double* bin, CgState* out, int slot, int spec,
           CgState* spec_out) {
  if (!c->have_op) return fail(c, MHD_EINVAL, "implicit solve before mhd_build_operator");
  CgArgs a = cg_args(c, bx, by);
  const long long maxit = (c->o.cg_max_iter > 0) ? c->o.cg_max_iter : 2LL * c->n;
  int rc0 = launch(c, 8, [&] { hipLaunchKernelGGL(k_cg_reset, dim3(1), dim3(1), 0, c->stream, c->st, maxit); });
  if (rc0) return rc0;
  const int nb = c->tiles;
  // outflow contexts run the CG kernels built for clamped neighbours
  const bool open = a.open_x | a.open_y;
  auto k_start_rhs = open ? k_cg_start_open<false> : k_cg_start<false>;
  auto k_start_b = open ? k_cg_start_open<true> : k_cg_start<true>;
  auto k_b0 = open ? k_cg_pass_b_open<true> : k_cg_pass_b<true>;
  auto k_a = open ? k_cg_pass_a_open : k_cg_pass_a;
  auto k_b = open ? k_cg_pass_b_open<false> : k_cg_pass_b<false>;
  int rc = bin ? launch_k(c, 2, k_start_rhs, dim3(nb), dim3(kTileX), a, bin)
               : launch_k(c, 2, k_start_b, dim3(nb), dim3(kTileX), a, bin);
  if (rc) return rc;
  rc = launch_k(c, 4, k_b0, dim3(nb), dim3(kTileX), a, (const double*)c->p[0], 0);
  if (rc) return rc;
  // the direction of iteration it lives in p[it & 1] (p[0] from the start kernel)
  long long it = 1;
  if (spec > 0) {
    // A(1) B(1) ... A(spec) B(spec) A(spec + 1): the last A takes the decision
    for (; it <= spec + 1; ++it) {
      const double* pold = c->p[(it - 1) & 1];
      double* pnew = c->p[it & 1];
      rc = launch_k(c, 3, k_a, dim3(nb), dim3(kTileX), a, pold, pnew, (int)it);
      if (rc) return rc;
      if (it == spec + 1) break;
      rc = launch_k(c, 4, k_b, dim3(nb), dim3(kTileX), a, (const double*)pnew, (int)it);
      if (rc) return rc;
    }
    if (spec_out) HIP_OR_FAIL(c, hipMemcpyAsync(spec_out, c->st, sizeof(CgState), hipMemcpyDeviceToDevice, c->stream));
    *out = CgState{};
    out->done = -1;
    out->spmv = -1;
    return MHD_OK;
  }
  for (;;) {
    const int chunk = (it == 1) ? (c->last_spmv[slot] > 0 ? c->last_spmv[slot] : 1) : 1;
    for (int k = 0; k < chunk; ++k, ++it) {
      const double* pold = c->p[(it - 1) & 1];
      double* pnew = c->p[it & 1];
      rc = launch_k(c, 3, k_a, dim3(nb), dim3(kTileX), a, pold, pnew, (int)it);
      if (rc) return rc;
      if (k + 1 == chunk) break;   // the chunk ends with an A: it holds the decision
      rc = launch_k(c, 4, k_b, dim3(nb), dim3(kTileX), a, (const double*)pnew, (int)it);
      if (rc) return rc;
    }
    HIP_OR_FAIL(c, hipMemcpyAsync(c->st_host, c->st, sizeof(CgState), hipMemcpyDeviceToHost, c->stream));
    HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
    if (c->st_host->done) break;
    // not converged: run B(it) for the A just launched, then continue
    const double* pv = c->p[it & 1];
    rc = launch_k(c, 4, k_b, dim3(nb), dim3(kTileX), a, pv, (int)it);
    if (rc) return rc;
    ++it;
    if (it > maxit + 4) return fail(c, MHD_EHIP, "CG state machine did not terminate");
  }
  *out = *c->st_host;
  if (out->spmv > 0) c->last_spmv[slot] = out->spmv;
  if (out->done == 2 && c->o.cg_fail_is_error) return fail(c, MHD_ECG, "CG reached the iteration limit");
  return MHD_OK;
}

double cg_error(const CgState& s) {
  if (s.done == 3 || s.done < 0) return 0.0;
  return std::sqrt(s.rn2 / s.rhs_norm2);
}

FinArgs fin_args(mhd_ctx* c, const double* bx, const double* by, double* bxo, double* byo, double coef) {
  FinArgs f{};
  f.nx = c->nx;
  f.ny = c->ny;
  f.open_x = c->o.bc_x == MHD_BC_OUTFLOW;
  f.open_y = c->o.bc_y == MHD_BC_OUTFLOW;
  f.dx = c->dx;
  f.dy = c->dy;
  f.inv_dx = 1.0 / c->dx;
  f.inv_dy = 1.0 / c->dy;
  f.eta = c->eta;
  f.c = coef;
  f.x = c->x;
  f.bx = bx;
  f.by = by;
  f.bx_out = bxo;
  f.by_out = byo;
  f.st = c->st;
  f.ty = c->fin_rows;
  f.tiles_y = c->fin_tiles_y;
  return f;
}

// fin: the tiles of the CFL partials (else those of the CG sweeps)
DiagArgs diag_args(mhd_ctx* c, bool fin) {
  DiagArgs d{};
  d.nx = c->nx;
  d.ny = c->ny;
  d.open_x = c->o.bc_x == MHD_BC_OUTFLOW;
  d.open_y = c->o.bc_y == MHD_BC_OUTFLOW;
  d.dx = c->dx;
  d.dy = c->dy;
  for (int q = 0; q < 6; ++q) d.s[q] = c->U[q];
  d.part = c->part;
  d.rho_raw = c->rho_raw;
  d.ty = fin ? c->fin_rows : c->tile_rows;
  d.tiles_y = fin ? c->fin_tiles_y : c->tiles_y;
  return d;
}

// Launches a per-tile diagnostic kernel and copies its partials to the host.
template <typename K>
int diag_partials(mhd_ctx* c, K kernel, int id, bool fin) {
  DiagArgs d = diag_args(c, fin);
  const int nb = fin ? c->fin_tiles : c->tiles;
  int rc = launch_k(c, id, kernel, dim3(nb), dim3(kTileX), d);
  if (rc) return rc;
  HIP_OR_FAIL(c, hipMemcpyAsync(c->part_host, c->part, 2 * sizeof(double) * nb, hipMemcpyDeviceToHost, c->stream));
  HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
  return MHD_OK;
}

// Host-side ghost fill of one padded array: periodic (Grid::fillOneField,
// Grid.hpp:143-161) or, on an outflow axis, the nearest inside value.
void fill_ghosts(double* f, int nx, int ny, int open_x, int open_y) {
  const int ng = 2, nxt = nx + 4, nyt = ny + 4;
  for (int j = ng; j < nyt - ng; ++j)
    for (int g = 0; g < ng; ++g) {
      f[j * nxt + g] = open_x ? f[j * nxt + ng] : f[j * nxt + nx + g];
      f[j * nxt + nx + ng + g] = open_x ? f[j * nxt + nx + ng - 1] : f[j * nxt + ng + g];
    }
  for (int i = 0; i < nxt; ++i)
    for (int g = 0; g < ng; ++g) {
      f[g * nxt + i] = open_y ? f[ng * nxt + i] : f[(ny + g) * nxt + i];
      f[(ny + ng + g) * nxt + i] = open_y ? f[(ny + ng - 1) * nxt + i] : f[(ng + g) * nxt + i];
    }
}

// Host transfers go through the context's pinned staging buffer: one
// contiguous DMA copy plus a row copy on the host (a strided copy from
// pageable memory runs at a small fraction of the link rate).
int to_host_padded(mhd_ctx* c, const double* dev, double* host, bool ghosts) {
  const size_t nx = (size_t)c->nx, nxt = nx + 4;
  HIP_OR_FAIL(c, hipMemcpyAsync(c->stage, dev, (size_t)c->n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
  for (int j = 0; j < c->ny; ++j) std::memcpy(host + (j + 2) * nxt + 2, c->stage + j * nx, nx * sizeof(double));
  if (ghosts) fill_ghosts(host, c->nx, c->ny, c->o.bc_x == MHD_BC_OUTFLOW, c->o.bc_y == MHD_BC_OUTFLOW);
  return MHD_OK;
}

int from_host_padded(mhd_ctx* c, const double* host, double* dev) {
  const size_t nx = (size_t)c->nx, nxt = nx + 4;
  HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));     // the staging buffer is free again
  for (int j = 0; j < c->ny; ++j) std::memcpy(c->stage + j * nx, host + (j + 2) * nxt + 2, nx * sizeof(double));
  HIP_OR_FAIL(c, hipMemcpyAsync(dev, c->stage, (size_t)c->n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  return MHD_OK;
}

int check_ctx(mhd_ctx* c) { return c ? MHD_OK : MHD_EINVAL; }

int reset_raw_density(mhd_ctx* c) {
  return launch(c, 8, [&] {
    hipLaunchKernelGGL(k_raw_reset, dim3(grid_1d(c->n)), dim3(256), 0, c->stream, (const double*)c->U[0], c->rho_raw,
                       (int)c->n);
  });
}

int copy_fields(mhd_ctx* c, double* const* dst, double* const* src, int n) {
  CopyArgs a{};
  for (int q = 0; q < n; ++q) {
    a.src[q] = src[q];
    a.dst[q] = dst[q];
  }
  a.nf = n;
  return launch(c, 8, [&] {
    hipLaunchKernelGGL(k_copy_fields, dim3(grid_1d(c->n)), dim3(256), 0, c->stream, a, (long long)c->n);
  });
}

int imex_step(mhd_ctx* c, double dt, mhd_step_stats* stats, StepRecord* rec) {
  mhd_step_stats st{};
  st.dt = dt;
  const Tableau& T = c->tab;
  int rc;
  if (rec) {
    if (rec->Un[0]) {
      rc = copy_fields(c, rec->Un, c->U, 6);
      if (rc) return rc;
    }
    if (!rec->use_given) rec->spmv[0] = rec->spmv[1] = 0;
  }
  // stage 1: f1 = L(U), S = U + dt A_ex[3] f1 (solvers.hpp:495-523),
  // W = U + dt A_ex[6] f1 (solvers.hpp:548-553) into F1
  StageArgs a = stage_args(c);
  for (int q = 0; q < 6; ++q) {
    a.s[q] = c->U[q];
    a.o1[q] = c->F1[q];
    a.o2[q] = c->S[q];
  }
  a.c_a = dt * T.A_ex3;
  a.c_b = dt * T.A_ex6;
  rc = run_stage(c, 1, a, 0);
  if (rc) return rc;
  const double diag_coeff = dt * T.A_im4;
  if (c->eta > 0.0) {
    // implicit 1: x = CG(eta J(S.B)), S.B += dt A_im[4] rate(x) (solvers.hpp:526-529)
    if (rec && rec->SsB[0]) {
      rc = copy_fields(c, rec->SsB, c->S + 4, 2);
      if (rc) return rc;
    }
    FinArgs f = fin_args(c, c->S[4], c->S[5], c->S[4], c->S[5], diag_coeff);
    if (rec && rec->use_given) {
      f.x = rec->x1;
      f.st = rec->st[0];
    } else {
      CgState cs;
      rc = cg_run(c, c->S[4], c->S[5], nullptr, &cs, 0, rec ? rec->spec[0] : 0, rec ? rec->st[0] : nullptr);
      if (rc) return rc;
      st.cg_iters[0] = cs.iters;
      st.cg_spmv[0] = cs.spmv;
      st.cg_error[0] = cg_error(cs);
      if (rec) {
        rec->spmv[0] = cs.spmv;
        if (cs.spmv != 0 && rec->x1) HIP_OR_FAIL(c, hipMemcpyAsync(rec->x1, c->x, (size_t)c->n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        if (rec->spec[0] == 0 && rec->st[0]) HIP_OR_FAIL(c, hipMemcpyAsync(rec->st[0], c->st, sizeof(CgState), hipMemcpyDeviceToDevice, c->stream));
      }
    }
    rc = launch_k(c, 5, k_finalize1, dim3(c->fin_tiles), dim3(kTileX), f);
    if (rc) return rc;
  }
  if (rec && rec->S[0]) {
    rc = copy_fields(c, rec->S, c->S, 6);
    if (rc) return rc;
  }
  // stage 2: U3* = W + dt A_ex[7] L(S) [+ dt A_im[7] g2] (solvers.hpp:535-566)
  StageArgs b = stage_args(c);
  for (int q = 0; q < 6; ++q) {
    b.s[q] = c->S[q];
    b.u[q] = c->F1[q];
    b.o1[q] = c->U[q];
  }
  b.c_b = dt * T.A_ex7;
  b.c_g = dt * T.A_im7;
  rc = run_stage(c, 2, b, 1);
  if (rc) return rc;
  if (c->eta > 0.0) {
    // implicit 2 and the final recombination (solvers.hpp:571-608)
    if (rec && rec->U3sB[0]) {
      rc = copy_fields(c, rec->U3sB, c->U + 4, 2);
      if (rc) return rc;
    }
    FinArgs f = fin_args(c, c->U[4], c->U[5], c->S[4], c->S[5], diag_coeff);
    if (rec && rec->use_given) {
      f.x = rec->x2;
      f.st = rec->st[1];
    } else {
      CgState cs;
      rc = cg_run(c, c->U[4], c->U[5], nullptr, &cs, 1, rec ? rec->spec[1] : 0, rec ? rec->st[1] : nullptr);
      if (rc) return rc;
      st.cg_iters[1] = cs.iters;
      st.cg_spmv[1] = cs.spmv;
      st.cg_error[1] = cg_error(cs);
      if (rec) {
        rec->spmv[1] = cs.spmv;
        if (cs.spmv != 0 && rec->x2) HIP_OR_FAIL(c, hipMemcpyAsync(rec->x2, c->x, (size_t)c->n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        if (rec->spec[1] == 0 && rec->st[1]) HIP_OR_FAIL(c, hipMemcpyAsync(rec->st[1], c->st, sizeof(CgState), hipMemcpyDeviceToDevice, c->stream));
      }
    }
    rc = launch_k(c, 6, k_finalize2, dim3(c->fin_tiles), dim3(kTileX), f);
    if (rc) return rc;
    std::swap(c->U[4], c->S[4]);
    std::swap(c->U[5], c->S[5]);
  }
  if (stats) *stats = st;
  return MHD_OK;
}

}  // namespace

extern "C" {

int mhd_default_opts(mhd_opts* o) {
  if (!o) return MHD_EINVAL;
  std::memset(o, 0, sizeof *o);
  o->gamma = 5.0 / 3.0;
  o->cfl = 0.45;
  o->cg_tol = DBL_EPSILON;
  o->cg_max_iter = 0;
  o->cg_fail_is_error = 0;
  o->march_rows = 0;
  o->bc_x = o->bc_y = MHD_BC_PERIODIC;
  return MHD_OK;
}

const char* mhd_last_error(const mhd_ctx* c) { return c ? c->err.c_str() : g_create_error.c_str(); }

int mhd_create(int nx, int ny, double dx, double dy, double eta, const mhd_opts* opts, mhd_ctx** out) {
  if (!out) return fail(nullptr, MHD_EINVAL, "out is NULL");
  *out = nullptr;
  if (nx < 4 || ny < 4) return fail(nullptr, MHD_EINVAL, "nx and ny must be >= 4");
  if ((long long)nx * ny > (1LL << 30)) return fail(nullptr, MHD_EINVAL, "nx * ny exceeds 2^30 cells");
  if (!(dx > 0) || !(dy > 0) || !std::isfinite(dx) || !std::isfinite(dy))
    return fail(nullptr, MHD_EINVAL, "dx and dy must be positive and finite");
  if (!(eta >= 0) || !std::isfinite(eta)) return fail(nullptr, MHD_EINVAL, "eta must be >= 0 and finite");
  mhd_opts o;
  mhd_default_opts(&o);
  if (opts) o = *opts;
  if (o.gamma != 5.0 / 3.0) return fail(nullptr, MHD_EINVAL, "only gamma = 5/3 is compiled in");
  if (!(o.cfl > 0) || !(o.cg_tol >= 0)) return fail(nullptr, MHD_EINVAL, "cfl must be > 0 and cg_tol >= 0");
  if (o.march_rows < 0) return fail(nullptr, MHD_EINVAL, "march_rows must be >= 0");
  if ((o.bc_x != MHD_BC_PERIODIC && o.bc_x != MHD_BC_OUTFLOW) || (o.bc_y != MHD_BC_PERIODIC && o.bc_y != MHD_BC_OUTFLOW))
    return fail(nullptr, MHD_EINVAL, "bc_x and bc_y must be MHD_BC_PERIODIC or MHD_BC_OUTFLOW");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(nullptr, MHD_EHIP, "no HIP device available");

  mhd_ctx* c = new mhd_ctx();
  c->nx = nx;
  c->ny = ny;
  c->n = (long long)nx * ny;
  c->dx = dx;
  c->dy = dy;
  c->eta = eta;
  c->o = o;
  auto bail = [&](const std::string& m, int code) {
    g_create_error = m;
    mhd_destroy(c);
    return code;
  };
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return bail("hipStreamCreate failed", MHD_EHIP);
  const size_t fb = (size_t)c->n * sizeof(double);
  for (int q = 0; q < 6; ++q) {
    if (hipMalloc(&c->U[q], fb) != hipSuccess || hipMalloc(&c->S[q], fb) != hipSuccess ||
        hipMalloc(&c->F1[q], fb) != hipSuccess)
      return bail("device allocation failed", MHD_ENOMEM);
  }
  // 2D sweeps: tiles of kTileX columns x tile_rows rows, about kTargetTiles of them
  {
    const int tiles_x = (nx + kTileX - 1) / kTileX;
    int want_y = kTargetTiles / tiles_x;
    if (want_y < 1) want_y = 1;
    int rows = (ny + want_y - 1) / want_y;
    if (rows < 4) rows = 4;
    c->tile_rows = rows;
    c->tiles_y = (ny + rows - 1) / rows;
    c->tiles = tiles_x * c->tiles_y;
    int fin_y = kFinTargetTiles / tiles_x;
    if (fin_y < 1) fin_y = 1;
    int fin_rows = (ny + fin_y - 1) / fin_y;
    if (fin_rows < 4) fin_rows = 4;
    c->fin_rows = fin_rows;
    c->fin_tiles_y = (ny + fin_rows - 1) / fin_rows;
    c->fin_tiles = tiles_x * c->fin_tiles_y;
  }
  const size_t pb = (size_t)c->tiles * 2 * sizeof(double);
  const size_t pd = (size_t)std::max(c->tiles, c->fin_tiles) * 2 * sizeof(double);   // part, part_host
  if (hipMalloc(&c->x, fb) != hipSuccess || hipMalloc(&c->r, fb) != hipSuccess || hipMalloc(&c->p[0], fb) != hipSuccess ||
      hipMalloc(&c->rho_raw, fb) != hipSuccess ||
      hipMalloc(&c->p[1], fb) != hipSuccess || hipMalloc(&c->part, pd) != hipSuccess ||
      hipMalloc(&c->part_s, pb * 2) != hipSuccess || hipMalloc(&c->part_a, pb) != hipSuccess ||
      hipMalloc(&c->part_b, pb) != hipSuccess ||
      hipMalloc(&c->st, sizeof(CgState)) != hipSuccess)
    return bail("device allocation failed", MHD_ENOMEM);
  if (hipHostMalloc(&c->st_host, sizeof(CgState)) != hipSuccess || hipHostMalloc(&c->part_host, pd) != hipSuccess ||
      hipHostMalloc(&c->stage, fb) != hipSuccess)
    return bail("pinned allocation failed", MHD_ENOMEM);
  for (int q = 0; q < 6; ++q) {
    if (hipMemsetAsync(c->U[q], 0, fb, c->stream) != hipSuccess) return bail("hipMemset failed", MHD_EHIP);
  if (hipMemsetAsync(c->rho_raw, 0, fb, c->stream) != hipSuccess) return bail("hipMemset failed", MHD_EHIP);
  }
  *c->st_host = CgState{};
  if (hipMemcpyAsync(c->st, c->st_host, sizeof(CgState), hipMemcpyHostToDevice, c->stream) != hipSuccess)
    return bail("device init failed", MHD_EHIP);
  // stage kernel: one wavefront per strip segment of stage_rows rows.  Pick
  // the segment length so that the launch fills the resident wavefront slots
  // of the device a whole number of times (a partial last round would idle
  // most of the chip), with segments of at least 32 rows; a grid large enough
  // to fill the slots with 32-row segments uses segments of at most 32 rows
  // (measured at 4096^2: 32-row segments in about three rounds run both
  // stages 6-7 % faster than 92-row segments in one round -- the waves running
  // at once then cover a narrower band of rows).
  if (o.march_rows > 0) {
    c->stage_rows = o.march_rows;
  } else {
    int dev = 0, ncu = 256, per_cu = 12;
    (void)hipGetDevice(&dev);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) == hipSuccess) ncu = prop.multiProcessorCount;
    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_stage<2>, 64, 0) == hipSuccess && occ > 0) per_cu = occ;
    const long long slots = (long long)ncu * per_cu;
    const long long strips = (nx + kStripOut - 1) / kStripOut;
    long long best_rows = ny, best_cost = -1;
    for (long long rounds = 1; rounds <= 64; ++rounds) {
      long long segs = (slots * rounds) / strips;
      if (segs < 1) continue;
      long long rows = (ny + segs - 1) / segs;
      if (rows < 32 && rounds > 1) break;
      if (rows > 32 && strips * ((ny + 31) / 32) >= slots) continue;
      long long waves = strips * ((ny + rows - 1) / rows);
      long long r_used = (waves + slots - 1) / slots;
      long long cost = r_used * (rows + 4);   // row steps of the slowest slot
      if (best_cost < 0 || cost < best_cost) { best_cost = cost; best_rows = rows; }
    }
    if (best_cost < 0) best_rows = ny < 32 ? ny : 32;   // no candidate left by the rules above
    if (best_rows < 4) {
      // a grid too small to fill the device: each wave runs alone on its
      // SIMD, bound by the dependency chain of its row steps, so the
      // shortest segments that still give each SIMD at most one wave win
      // (measured, stage kernel per launch: 128^2 one row 9.5 us against 14.4
      // at four, 256^2 two rows 11.7 against 14.5); otherwise four rows
      const long long simds = (long long)ncu * 4;
      long long rows = 1;
      while (rows < 4 && strips * ((ny + rows - 1) / rows) > simds) ++rows;
      best_rows = rows;
    }
    c->stage_rows = (int)best_rows;
  }
  if (hipStreamSynchronize(c->stream) != hipSuccess) return bail("device init failed", MHD_EHIP);
  *out = c;
  return MHD_OK;
}

int mhd_destroy(mhd_ctx* c) {
  if (!c) return MHD_EINVAL;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  destroy_adjoint(c);
  destroy_resident(c);
  for (int k = 0; k < kNumKernelIds; ++k) {
    for (auto& pr : c->pending[k]) {
      (void)hipEventDestroy(pr.first);
      (void)hipEventDestroy(pr.second);
    }
  }
  for (int q = 0; q < 6; ++q) {
    if (c->U[q]) (void)hipFree(c->U[q]);
    if (c->S[q]) (void)hipFree(c->S[q]);
    if (c->F1[q]) (void)hipFree(c->F1[q]);
  }
  double* bufs[] = {c->x, c->r, c->p[0], c->p[1], c->part, c->part_s, c->part_a, c->part_b, c->rho_raw};
  for (double* b : bufs)
    if (b) (void)hipFree(b);
  if (c->st) (void)hipFree(c->st);
  if (c->st_host) (void)hipHostFree(c->st_host);
  if (c->part_host) (void)hipHostFree(c->part_host);
  if (c->stage) (void)hipHostFree(c->stage);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return MHD_OK;
}

int mhd_upload(mhd_ctx* c, const double* rho, const double* mx, const double* my, const double* E,
               const double* bxf, const double* byf) {
  if (check_ctx(c)) return MHD_EINVAL;
  const double* src[6] = {rho, mx, my, E, bxf, byf};
  for (int q = 0; q < 6; ++q)
    if (!src[q]) return fail(c, MHD_EINVAL, "mhd_upload: NULL field");
  for (int q = 0; q < 6; ++q) {
    int rc0 = from_host_padded(c, src[q], c->U[q]);
    if (rc0) return rc0;
  }
  int rc = launch(c, 8, [&] {
    hipLaunchKernelGGL(k_floor_rho_raw, dim3(grid_1d(c->n)), dim3(256), 0, c->stream, c->U[0], c->rho_raw, (int)c->n);
  });
  if (rc) return rc;
  HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
  return MHD_OK;
}

int mhd_download(mhd_ctx* c, double* rho, double* mx, double* my, double* E, double* bxf, double* byf, double* vx,
                 double* vy, double* p, double* bxc, double* byc) {
  if (check_ctx(c)) return MHD_EINVAL;
  double* dst[6] = {rho, mx, my, E, bxf, byf};
  for (int q = 0; q < 6; ++q)
    if (dst[q]) {
      int rc = to_host_padded(c, c->U[q], dst[q], true);
      if (rc) return rc;
    }
  double* pd[5] = {vx, vy, p, bxc, byc};
  bool any = false;
  for (double* q : pd) any = any || q;
  if (any) {
    DiagArgs d = diag_args(c, false);
    int rc = launch(c, 8, [&] {
      hipLaunchKernelGGL(k_prims, dim3(grid_1d(c->n)), dim3(256), 0, c->stream, d, c->F1[0], c->F1[1], c->F1[2], c->F1[3],
                         c->F1[4]);
    });
    if (rc) return rc;
    for (int q = 0; q < 5; ++q)
      if (pd[q]) {
        rc = to_host_padded(c, c->F1[q], pd[q], true);
        if (rc) return rc;
      }
  }
  return MHD_OK;
}

int mhd_upload_device(mhd_ctx* c, const double* const* fields) {
  if (check_ctx(c)) return MHD_EINVAL;
  if (!fields) return fail(c, MHD_EINVAL, "mhd_upload_device: NULL field array");
  double* src[6];
  for (int q = 0; q < 6; ++q) {
    if (!fields[q]) return fail(c, MHD_EINVAL, "mhd_upload_device: NULL field");
    src[q] = const_cast<double*>(fields[q]);
  }
  int rc = copy_fields(c, c->U, src, 6);
  if (rc) return rc;
  rc = launch(c, 8, [&] {
    hipLaunchKernelGGL(k_floor_rho_raw, dim3(grid_1d(c->n)), dim3(256), 0, c->stream, c->U[0], c->rho_raw, (int)c->n);
  });
  if (rc) return rc;
  HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
  return MHD_OK;
}

int mhd_download_device(mhd_ctx* c, double* const* fields) {
  if (check_ctx(c)) return MHD_EINVAL;
  if (!fields) return fail(c, MHD_EINVAL, "mhd_download_device: NULL field array");
  for (int q = 0; q < 6; ++q)
    if (fields[q]) {
      double* dst[1] = {fields[q]};
      int rc = copy_fields(c, dst, &c->U[q], 1);
      if (rc) return rc;
    }
  HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
  return MHD_OK;
}

static int max_speed(mhd_ctx* c, double* maxs, double* nbad) {
  int rc = diag_partials(c, k_cfl, 7, true);
  if (rc) return rc;
  // exact: a max (NaN skipped as the reference's `>` skips it) and a count
  double m = 1e-8, bad = 0.0;
  for (int b = 0; b < c->fin_tiles; ++b) {
    const double v = c->part_host[2 * b];
    if (v > m) m = v;
    bad += c->part_host[2 * b + 1];
  }
  *maxs = m;
  *nbad = bad;
  return MHD_OK;
}

int mhd_cfl_dt(mhd_ctx* c, double* dt) {
  if (check_ctx(c) || !dt) return fail(c, MHD_EINVAL, "mhd_cfl_dt: bad argument");
  double maxs = 0, nbad = 0;
  int rc = max_speed(c, &maxs, &nbad);
  if (rc) return rc;
  // CFL_condition_checker, diagnostics.hpp:52-86, with the Courant number opts.cfl
  const double h = (c->dy < c->dx) ? c->dy : c->dx;
  double dt_res = INFINITY;
  if (c->eta != 0.0) dt_res = h * h / (4.0 * c->eta);
  const double courant = c->o.cfl / 0.9;
  double dt_hyp = courant * h / maxs;
  double lim = (dt_res < dt_hyp) ? dt_res : dt_hyp;
  *dt = lim * 0.9;
  if (nbad > 0) return fail(c, MHD_ENONFINITE, "non-finite wave speed in " + std::to_string((long long)nbad) + " cells");
  return MHD_OK;
}

int mhd_build_operator(mhd_ctx* c, double alpha) {
  if (check_ctx(c)) return MHD_EINVAL;
  if (!std::isfinite(alpha) || alpha < 0) return fail(c, MHD_EINVAL, "alpha must be finite and >= 0");
  c->alpha = alpha;
  c->have_op = true;
  return MHD_OK;
}

int mhd_imex_step(mhd_ctx* c, double dt, mhd_step_stats* stats) {
  if (check_ctx(c)) return MHD_EINVAL;
  if (!std::isfinite(dt)) return fail(c, MHD_EINVAL, "dt must be finite");
  return imex_step(c, dt, stats, nullptr);
}

int mhd_explicit_rhs(mhd_ctx* c, double* drho, double* dmx, double* dmy, double* dE, double* dbxf, double* dbyf) {
  if (check_ctx(c)) return MHD_EINVAL;
  double* dst[6] = {drho, dmx, dmy, dE, dbxf, dbyf};
  StageArgs a = stage_args(c);
  for (int q = 0; q < 6; ++q) {
    a.s[q] = c->U[q];
    a.o1[q] = c->F1[q];
  }
  int rc = run_stage(c, 0, a, 8);
  if (rc) return rc;
  for (int q = 0; q < 6; ++q)
    if (dst[q]) {
      rc = to_host_padded(c, c->F1[q], dst[q], false);
      if (rc) return rc;
    }
  return MHD_OK;
}

int mhd_resistive_term(mhd_ctx* c, int implicit, double* bx_rate, double* by_rate, int* iters) {
  if (check_ctx(c) || !bx_rate || !by_rate) return fail(c, MHD_EINVAL, "mhd_resistive_term: bad argument");
  int rc;
  FinArgs f = fin_args(c, c->U[4], c->U[5], c->F1[0], c->F1[1], 0.0);
  int zero = 0;
  if (implicit) {
    CgState cs;
    rc = cg_run(c, c->U[4], c->U[5], nullptr, &cs, 0);
    if (rc) return rc;
    if (iters) *iters = cs.iters;
    zero = cs.spmv == 0;
  } else {
    CgArgs a = cg_args(c, c->U[4], c->U[5]);
    rc = launch(c, 8, [&] { hipLaunchKernelGGL(k_cg_rhs, dim3(c->tiles), dim3(kTileX), 0, c->stream, a); });
    if (rc) return rc;
    f.x = c->r;
    if (iters) *iters = 0;
  }
  rc = launch(c, 8, [&] { hipLaunchKernelGGL(k_rate, dim3(grid_1d(c->n)), dim3(256), 0, c->stream, f, zero); });
  if (rc) return rc;
  rc = to_host_padded(c, c->F1[0], bx_rate, false);
  if (rc) return rc;
  return to_host_padded(c, c->F1[1], by_rate, false);
}

int mhd_cg_solve(mhd_ctx* c, const double* b, double* x, int* iters, double* error) {
  if (check_ctx(c) || !b || !x) return fail(c, MHD_EINVAL, "mhd_cg_solve: bad argument");
  const size_t fb = (size_t)c->n * sizeof(double);
  HIP_OR_FAIL(c, hipMemcpyAsync(c->F1[0], b, fb, hipMemcpyHostToDevice, c->stream));
  CgState cs;
  int rc = cg_run(c, c->U[4], c->U[5], c->F1[0], &cs, 0);
  if (rc) return rc;
  if (cs.spmv == 0) {
    std::memset(x, 0, fb);
  } else {
    HIP_OR_FAIL(c, hipMemcpyAsync(x, c->x, fb, hipMemcpyDeviceToHost, c->stream));
    HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
  }
  if (iters) *iters = cs.iters;
  if (error) *error = cg_error(cs);
  return MHD_OK;
}

int mhd_divb(mhd_ctx* c, double* rms, double* maxabs) {
  if (check_ctx(c) || !rms || !maxabs) return fail(c, MHD_EINVAL, "mhd_divb: bad argument");
  int rc = diag_partials(c, k_divb, 8, false);
  if (rc) return rc;
  double sum = 0.0, mx = 0.0;
  for (int b = 0; b < c->tiles; ++b) {      // fixed order: deterministic
    sum += c->part_host[2 * b];
    if (c->part_host[2 * b + 1] > mx) mx = c->part_host[2 * b + 1];
  }
  *rms = std::sqrt(sum / (double)(c->nx * c->ny));
  *maxabs = mx;
  return MHD_OK;
}

int mhd_run(mhd_ctx* c, double t_end, double dt_override, int max_steps, int* steps_out, double* t_out,
            double* step_dt, int* cg_iters) {
  if (check_ctx(c)) return MHD_EINVAL;
  if (!std::isfinite(t_end) || !std::isfinite(dt_override) || dt_override < 0)
    return fail(c, MHD_EINVAL, "mhd_run: bad t_end or dt_override");
  const double A_im4 = c->tab.A_im4;
  double opt;
  int rc = mhd_cfl_dt(c, &opt);                                   // integrator_main.cpp:55
  if (rc) return rc;
  double dt0 = (dt_override == 0.0) ? opt : dt_override;
  double mc = dt0 * c->eta;
  rc = mhd_build_operator(c, mc * A_im4);                         // :58-60
  if (rc) return rc;
  // :67: primitives of the floored state, which the first CFL step reads
  rc = reset_raw_density(c);
  if (rc) return rc;
  double prev = dt0, t = 0.0;
  int steps = 0;
  while (t < t_end && (max_steps < 0 || steps < max_steps)) {     // :207
    // :211-212; with an override the CFL value only feeds the reference's warning print
    double dt = dt_override;
    if (dt_override == 0.0) {
      rc = mhd_cfl_dt(c, &dt);
      if (rc) return rc;
    }
    if (t + dt > t_end) dt = t_end - t;                           // :214
    if (c->eta > 0.0) {
      if (std::fabs(dt - prev) > 1e-10 * dt) {                     // :219
        double m = dt * c->eta;
        rc = mhd_build_operator(c, m * A_im4);
        if (rc) return rc;
      }
    }
    prev = dt;
    mhd_step_stats st;
    rc = mhd_imex_step(c, dt, &st);                               // :230
    if (rc) return rc;
    t += dt;
    if (step_dt) step_dt[steps] = dt;
    if (cg_iters) {
      cg_iters[2 * steps] = st.cg_iters[0];
      cg_iters[2 * steps + 1] = st.cg_iters[1];
    }
    steps++;
  }
  if (steps_out) *steps_out = steps;
  if (t_out) *t_out = t;
  return MHD_OK;
}

int mhd_profile_enable(mhd_ctx* c, int mask) {
  if (check_ctx(c)) return MHD_EINVAL;
  int rc = flush_profile(c);
  if (rc) return rc;
  c->prof_mask = (unsigned)mask & ((1u << kNumKernelIds) - 1u);
  for (int k = 0; k < kNumKernelIds; ++k) {
    c->prof_n[k] = 0;
    c->prof_ms[k] = 0.0;
  }
  return MHD_OK;
}

int mhd_profile_get(mhd_ctx* c, int id, long long* launches, double* total_ms) {
  if (check_ctx(c) || id < 0 || id >= kNumKernelIds) return fail(c, MHD_EINVAL, "mhd_profile_get: bad argument");
  int rc = flush_profile(c);
  if (rc) return rc;
  if (launches) *launches = c->prof_n[id];
  if (total_ms) *total_ms = c->prof_ms[id];
  return MHD_OK;
}

int mhd_calibration_copy(mhd_ctx* c, long long nbytes) {
  if (check_ctx(c) || nbytes <= 0) return fail(c, MHD_EINVAL, "mhd_calibration_copy: bad argument");
  long long n = nbytes / (long long)sizeof(double);
  double *src = nullptr, *dst = nullptr;
  HIP_OR_FAIL(c, hipMalloc(&src, n * sizeof(double)));
  HIP_OR_FAIL(c, hipMalloc(&dst, n * sizeof(double)));
  HIP_OR_FAIL(c, hipMemsetAsync(src, 0, n * sizeof(double), c->stream));
  int rc = launch(c, 8, [&] { hipLaunchKernelGGL(k_copy, dim3(grid_1d(n)), dim3(256), 0, c->stream, (const double*)src, dst, n); });
  HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
  (void)hipFree(src);
  (void)hipFree(dst);
  return rc;
}

int mhd_calibration_mix(mhd_ctx* c, int n_in, int n_out, long long n, int reps, double* ms_per_launch) {
  if (check_ctx(c) || n <= 0 || reps <= 0 || !ms_per_launch)
    return fail(c, MHD_EINVAL, "mhd_calibration_mix: bad argument");
  void (*k)(MixArgs, long long) = nullptr;
  const int mix = n_in * 100 + n_out;
  switch (mix) {
    case 101: k = k_stream_mix<1, 1>; break;
    case 201: k = k_stream_mix<2, 1>; break;
    case 202: k = k_stream_mix<2, 2>; break;
    case 302: k = k_stream_mix<3, 2>; break;
    case 702: k = k_stream_mix<7, 2>; break;
    case 612: k = k_stream_mix<6, 12>; break;
    case 1206: k = k_stream_mix<12, 6>; break;
    default:
      return fail(c, MHD_EINVAL, "mhd_calibration_mix: mixes (in, out): (1,1) (2,1) (2,2) (3,2) (7,2) (6,12) (12,6)");
  }
  const int na = n_in + n_out;
  std::vector<double*> arr(na, nullptr);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = MHD_OK;
  auto release = [&] {
    for (double* p : arr)
      if (p) (void)hipFree(p);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  };
  for (int q = 0; q < na; ++q) {
    if (hipMalloc(&arr[q], (size_t)n * sizeof(double)) != hipSuccess) {
      arr[q] = nullptr;
      release();
      return fail(c, MHD_ENOMEM, "mhd_calibration_mix: allocation failed");
    }
  }
  MixArgs a{};
  for (int q = 0; q < n_in; ++q) a.in[q] = arr[q];
  for (int q = 0; q < n_out; ++q) a.out[q] = arr[n_in + q];
  hipError_t e = hipSuccess;
  for (int q = 0; q < n_in && e == hipSuccess; ++q) e = hipMemsetAsync(arr[q], 0, (size_t)n * sizeof(double), c->stream);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  const dim3 grid((unsigned)std::min<long long>((n + 255) / 256, 65536LL));
  for (int w = 0; w < 2 && e == hipSuccess; ++w) {
    hipLaunchKernelGGL(k, grid, dim3(256), 0, c->stream, a, n);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipEventRecord(e0, c->stream);
  for (int r = 0; r < reps && e == hipSuccess; ++r) {
    hipLaunchKernelGGL(k, grid, dim3(256), 0, c->stream, a, n);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipEventRecord(e1, c->stream);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float ms = 0.0f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  if (e != hipSuccess) rc = fail(c, MHD_EHIP, std::string("mhd_calibration_mix: ") + hipGetErrorString(e));
  else *ms_per_launch = (double)ms / reps;
  release();
  return rc;
}

int mhd_synchronize(mhd_ctx* c) {
  if (check_ctx(c)) return MHD_EINVAL;
  HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
  return MHD_OK;
}

}  // extern "C"
