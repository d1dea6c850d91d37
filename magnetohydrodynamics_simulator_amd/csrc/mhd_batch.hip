// mhd_batch.hip -- C ABI of batched ensembles (include/mhd_mi355x.h, mhd_batch_*).
//
// Many independent periodic runs of different sizes, resistivities and time
// steps advance together: every field of every run lives in one pooled array
// per field (run r at cells [base_r, base_r + nx_r ny_r)), and each kernel of
// the step is launched once for all runs with a job table that maps blocks to
// (run, block within the run).  The per-run driver (CFL dt, clamp to t_end,
// operator rebuild, t and step count) runs on the device (kb_dt), so a step
// needs no host round trip except the convergence check of each CG solve,
// which is one for the whole batch.  The sequence of a step is the single-run
// one (mhd_capi.hip):
//   kb_cfl, kb_dt, kb_stage<1>, CG (slot 0), kb_finalize1, kb_stage<2>,
//   CG (slot 1), kb_finalize2
// and each run's arithmetic is that of the single-run kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "../../include/mhd_mi355x.h"
#include "mhd_batch.hpp"
#include "mhd_kernels.hpp"

using namespace mhd;

namespace mhd_batch_host {

thread_local std::string g_batch_create_error;

int bfail(mhd_batch* b, int code, const std::string& msg) {
  if (b) b->err = msg;
  else g_batch_create_error = msg;
  return code;
}

}  // namespace mhd_batch_host

using namespace mhd_batch_host;

namespace {

// Candidate launch shapes of the autotuner: rows per stage-kernel segment and
// rows per 2D tile.  Buffers are sized for the smallest of each.
constexpr int kStageRowChoices[] = {16, 32, 64};
constexpr int kTileRowChoices[] = {16, 32, 64};

// Lays out the runs: pooled offsets, stage segments and 2D tiles, job tables.
void plan(mhd_batch* b, std::vector<int2>& stage_jobs, std::vector<int2>& tile_jobs) {
  long long base = 0;
  int part_base = 0;
  stage_jobs.clear();
  tile_jobs.clear();
  for (int r = 0; r < b->nruns; ++r) {
    RunGeom& g = b->geom_h[r];
    g.base = base;
    base += (long long)g.nx * g.ny;
    g.rows = std::min(b->stage_rows, g.ny);
    g.strips = (g.nx + kStripOut - 1) / kStripOut;
    const int segs = (g.ny + g.rows - 1) / g.rows;
    for (int w = 0; w < g.strips * segs; ++w) stage_jobs.push_back(make_int2(r, w));
    g.ty = std::max(4, std::min(b->tile_rows, g.ny));
    g.tiles_y = (g.ny + g.ty - 1) / g.ty;
    const int tiles_x = (g.nx + kTileX - 1) / kTileX;
    g.tiles = tiles_x * g.tiles_y;
    g.part_base = part_base;
    part_base += g.tiles;
    for (int t = 0; t < g.tiles; ++t) tile_jobs.push_back(make_int2(r, t));
  }
  b->total = base;
  b->total_tiles = part_base;
}

// Lays the runs out with the given launch shape and uploads geometry and job
// tables (the schedule fields of the geometry are kept).
void drop_graphs(mhd_batch* b) {
  for (BatchGraph& g : b->graphs) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    g = BatchGraph{};
  }
}

int replan(mhd_batch* b, int stage_rows, int tile_rows) {
  b->cfl_fresh = false;
  drop_graphs(b);        // captured launches carry the old grid sizes
  b->stage_rows = stage_rows;
  b->tile_rows = tile_rows;
  std::vector<int2> sj, tj;
  plan(b, sj, tj);
  if (sj.size() > b->cap_stage_jobs || tj.size() > b->cap_tiles)
    return bfail(b, MHD_EINVAL, "launch shape exceeds the planned capacity");
  b->n_stage_jobs = (int)sj.size();
  b->n_tile_jobs = (int)tj.size();
  BHIP(b, hipMemcpyAsync(b->geom, b->geom_h.data(), sizeof(RunGeom) * b->nruns, hipMemcpyHostToDevice, b->stream));
  BHIP(b, hipMemcpyAsync(b->stage_jobs, sj.data(), sizeof(int2) * sj.size(), hipMemcpyHostToDevice, b->stream));
  BHIP(b, hipMemcpyAsync(b->tile_jobs, tj.data(), sizeof(int2) * tj.size(), hipMemcpyHostToDevice, b->stream));
  BHIP(b, hipStreamSynchronize(b->stream));
  return MHD_OK;
}

BatchCgArgs cg_args(mhd_batch* b, const CgCall& c) {
  BatchCgArgs a{};
  a.geom = b->geom;
  a.rs = c.rs;
  a.jobs = b->tile_jobs;
  a.bx = c.bx;
  a.by = c.by;
  a.r = b->r;
  a.p0 = b->p[0];
  a.x = c.x;
  a.part_s = b->part_s;
  a.part_a = b->part_a;
  a.part_b = b->part_b;
  a.st = c.st;
  a.tol = b->o.cg_tol;
  return a;
}

}  // namespace

namespace mhd_batch_host {

// The single-run kernel sequence of a solve (cg_run in mhd_capi.hip) over
// all runs at once.
int cg_batch(mhd_batch* b, const CgCall& c, int* most_spmv) {
  BatchCgArgs a = cg_args(b, c);
  const int nb = b->n_tile_jobs;
  if (c.bin) BLAUNCH(b, kb_start_rhs(b), dim3(nb), dim3(kTileX), 0, b->stream, a, c.bin);
  else BLAUNCH(b, kb_start(b), dim3(nb), dim3(kTileX), 0, b->stream, a);
  BLAUNCH(b, kb_pass_b<true>(b), dim3(nb), dim3(kTileX), 0, b->stream, a, (const double*)b->p[0], 0);
  long long it = 1;
  if (c.spec > 0) {
    // A(1) B(1) ... A(spec) B(spec) A(spec + 1): the last A takes the decision
    for (; it <= c.spec + 1; ++it) {
      const double* pold = b->p[(it - 1) & 1];
      double* pnew = b->p[it & 1];
      BLAUNCH(b, kb_pass_a(b), dim3(nb), dim3(kTileX), 0, b->stream, a, pold, pnew, (int)it);
      if (it == c.spec + 1) break;
      BLAUNCH(b, kb_pass_b<false>(b), dim3(nb), dim3(kTileX), 0, b->stream, a, (const double*)pnew, (int)it);
    }
    return MHD_OK;
  }
  long long maxit = 0;
  for (const RunGeom& g : b->geom_h) maxit = std::max(maxit, g.max_iters);
  for (;;) {
    const int chunk = (it == 1) ? std::max(1, b->last_spmv[c.slot]) : 1;
    for (int k = 0; k < chunk; ++k, ++it) {
      const double* pold = b->p[(it - 1) & 1];
      double* pnew = b->p[it & 1];
      BLAUNCH(b, kb_pass_a(b), dim3(nb), dim3(kTileX), 0, b->stream, a, pold, pnew, (int)it);
      if (k + 1 == chunk) break;
      BLAUNCH(b, kb_pass_b<false>(b), dim3(nb), dim3(kTileX), 0, b->stream, a, (const double*)pnew, (int)it);
    }
    BHIP(b, hipMemcpyAsync(b->st_host, c.st, sizeof(CgState) * b->nruns, hipMemcpyDeviceToHost, b->stream));
    BHIP(b, hipStreamSynchronize(b->stream));
    bool all = true;
    for (int r = 0; r < b->nruns; ++r) all = all && b->st_host[r].done;
    if (all) break;
    const double* pv = b->p[it & 1];
    BLAUNCH(b, kb_pass_b<false>(b), dim3(nb), dim3(kTileX), 0, b->stream, a, pv, (int)it);
    ++it;
    if (it > maxit + 4) return bfail(b, MHD_EHIP, "CG state machine did not terminate");
  }
  int most = 0;
  for (int r = 0; r < b->nruns; ++r) {
    const CgState& s = b->st_host[r];
    if (s.done != 3) most = std::max(most, s.spmv);
    // runs that stepped (rs_host arrived before the first state copy)
    if (c.slot < 2 && b->rs_host[r].active) b->cg_iters_last[2 * r + c.slot] = s.iters;
    if (s.done == 2 && b->o.cg_fail_is_error)
      return bfail(b, MHD_ECG, "run " + std::to_string(r) + ": CG reached the iteration limit");
  }
  if (most > 0) b->last_spmv[c.slot] = most;
  if (most_spmv) *most_spmv = most;
  return MHD_OK;
}

// Ghosts of a padded (ny+4) x (nx+4) field: periodic (Grid::fillOneField,
// Grid.hpp:143-161) or, on an outflow axis, the nearest inside value.
void fill_ghosts_padded(double* f, int nx, int ny, int open_x, int open_y) {
  const size_t nxt = (size_t)nx + 4;
  for (int j = 2; j < ny + 2; ++j)
    for (int k = 0; k < 2; ++k) {
      f[j * nxt + k] = open_x ? f[j * nxt + 2] : f[j * nxt + nx + k];
      f[j * nxt + nx + 2 + k] = open_x ? f[j * nxt + nx + 1] : f[j * nxt + 2 + k];
    }
  for (size_t i = 0; i < nxt; ++i)
    for (int k = 0; k < 2; ++k) {
      f[k * nxt + i] = open_y ? f[2 * nxt + i] : f[(ny + k) * nxt + i];
      f[(ny + 2 + k) * nxt + i] = open_y ? f[(ny + 1) * nxt + i] : f[(2 + k) * nxt + i];
    }
}

// Interior of a padded host field of run `run` into the pooled device array,
// through the pinned staging buffer (a strided copy from pageable memory is slow).
int stage_in(mhd_batch* b, int run, const double* host, double* pooled) {
  const RunGeom& g = b->geom_h[run];
  const size_t nx = (size_t)g.nx, nxt = nx + 4;
  BHIP(b, hipStreamSynchronize(b->stream));   // the staging buffer is free again
  for (int j = 0; j < g.ny; ++j) std::memcpy(b->stage + j * nx, host + (j + 2) * nxt + 2, nx * sizeof(double));
  BHIP(b, hipMemcpyAsync(pooled + g.base, b->stage, nx * g.ny * sizeof(double), hipMemcpyHostToDevice, b->stream));
  return MHD_OK;
}

// Run `run` of a pooled device array into a padded host field, ghosts filled.
int stage_out(mhd_batch* b, int run, const double* pooled, double* host) {
  const RunGeom& g = b->geom_h[run];
  const size_t nx = (size_t)g.nx, nxt = nx + 4;
  BHIP(b, hipMemcpyAsync(b->stage, pooled + g.base, nx * g.ny * sizeof(double), hipMemcpyDeviceToHost, b->stream));
  BHIP(b, hipStreamSynchronize(b->stream));
  for (int j = 0; j < g.ny; ++j) std::memcpy(host + (j + 2) * nxt + 2, b->stage + (size_t)j * nx, nx * sizeof(double));
  fill_ghosts_padded(host, g.nx, g.ny, g.open & 1, (g.open >> 1) & 1);
  return MHD_OK;
}

BatchFinArgs fin_args(mhd_batch* b, const RunState* rs, const double* x, const CgState* st, const double* bx,
                      const double* by, double* bxo, double* byo) {
  BatchFinArgs f{};
  f.geom = b->geom;
  f.rs = rs;
  f.jobs = b->tile_jobs;
  f.x = x;
  f.bx = bx;
  f.by = by;
  f.bx_out = bxo;
  f.by_out = byo;
  f.st = st;
  return f;
}

}  // namespace mhd_batch_host

namespace {

int check_run(mhd_batch* b, int run) {
  if (!b) return MHD_EINVAL;
  if (run < 0 || run >= b->nruns) return bfail(b, MHD_EINVAL, "run index out of range");
  return MHD_OK;
}

// ---------------------------------------------------------------------------
// Device-resident groups of batch steps (mhd_batch_run in group mode).
//
// The host enqueues `group` steps of the whole batch and reads the driver
// state back once per group.  Every CG solve gets a launched-iteration
// budget (the longest solve of that slot in the last group, plus a margin);
// a one-block check after each solve halts the group (control word
// kCtlHalt) if a stepping run's solve is unfinished, and every later batched
// kernel then returns at once.  At the read-back the host continues that
// solve with host-checked iterations, finishes the step and enqueues the
// group's remaining steps.  The kernels and their order are those of
// mhd_batch_step, so the states, time steps and iteration counts are
// identical to stepping one host-checked step at a time.
// ---------------------------------------------------------------------------

struct BLoop {
  double* B0[2];     // the state's face B at the start of the call (U[4], U[5])
  double* B1[2];     // the other pair (S[4], S[5])
  int spec[2];       // launched iterations per solve: A(1) B(1) ... A(spec) B(spec) A(spec + 1)
};

// The B pointers after `completed` steps of the call (every step of a batch
// with a resistive run ends with the swap of the final recombination).
void set_bparity(mhd_batch* b, const BLoop& L, long long completed) {
  const bool odd = b->any_eta && (completed & 1);
  for (int k = 0; k < 2; ++k) {
    b->U[4 + k] = odd ? L.B1[k] : L.B0[k];
    b->S[4 + k] = odd ? L.B0[k] : L.B1[k];
  }
}

BatchDiagArgs group_diag(mhd_batch* b) {
  BatchDiagArgs d{};
  d.geom = b->geom;
  d.rs = b->rs;
  d.jobs = b->tile_jobs;
  for (int q = 0; q < 6; ++q) d.s[q] = b->U[q];
  d.part = b->part;
  d.ctl = b->ctl;
  d.rho_raw = b->rho_raw;
  return d;
}

BatchCgArgs group_cg(mhd_batch* b, int slot) {
  CgCall cc;
  cc.slot = slot;
  cc.bx = slot == 0 ? b->S[4] : b->U[4];
  cc.by = slot == 0 ? b->S[5] : b->U[5];
  cc.x = b->x;
  cc.st = b->st[slot];
  cc.rs = b->rs;
  BatchCgArgs a = cg_args(b, cc);
  a.ctl = b->ctl;
  return a;
}

// Launched iterations of a solve: start, B(0), then A(1) B(1) ... A(spec) B(spec)
// A(spec + 1) (the last A takes the convergence decision), then the check.
int enqueue_group_cg(mhd_batch* b, int slot, int spec) {
  const BatchCgArgs a = group_cg(b, slot);
  const int nb = b->n_tile_jobs;
  BLAUNCH(b, kb_start(b), dim3(nb), dim3(kTileX), 0, b->stream, a);
  BLAUNCH(b, kb_pass_b<true>(b), dim3(nb), dim3(kTileX), 0, b->stream, a, (const double*)b->p[0], 0);
  for (int it = 1; it <= spec + 1; ++it) {
    BLAUNCH(b, kb_pass_a(b), dim3(nb), dim3(kTileX), 0, b->stream, a, (const double*)b->p[(it - 1) & 1], b->p[it & 1],
            it);
    if (it == spec + 1) break;
    BLAUNCH(b, kb_pass_b<false>(b), dim3(nb), dim3(kTileX), 0, b->stream, a, (const double*)b->p[it & 1], it);
  }
  BLAUNCH(b, kb_check_solve, dim3(1), dim3(256), 0, b->stream, (const RunState*)b->rs, (const CgState*)b->st[slot],
          b->nruns, slot, b->o.cg_fail_is_error, b->ctl, b->iters);
  return MHD_OK;
}

// Phases of a step: 0 CFL partials (unless the last recombination formed
// them) and kb_dt, 1 stage 1, 2 solve 1 and its check, 3 finaliser 1,
// 4 stage 2, 5 solve 2 and its check, 6 final recombination (+ the B swap).
int enqueue_group_step(mhd_batch* b, BLoop& L, int from) {
  const BatchTableau& T = b->tab;
  const int nt = b->n_tile_jobs;
  const BatchDiagArgs d = group_diag(b);
  const bool fused = b->any_eta && b->fused_cfl;
  if (from <= 0) {
    if (!fused) BLAUNCH(b, kb_cfl, dim3(nt), dim3(kTileX), 0, b->stream, d);
    BLAUNCH(b, kb_dt, dim3(b->nruns), dim3(256), 0, b->stream, (const RunGeom*)b->geom, b->rs, (const double*)b->part,
            b->st[0], b->st[1], b->nruns, step_factors(T), 1, b->ctl);
  }
  if (from <= 1) {
    BatchStageArgs s1{};
    s1.geom = b->geom;
    s1.rs = b->rs;
    s1.jobs = b->stage_jobs;
    s1.ctl = b->ctl;
    for (int q = 0; q < 6; ++q) {
      s1.s[q] = b->U[q];
      s1.o1[q] = b->W[q];
      s1.o2[q] = b->S[q];
    }
    BLAUNCH(b, kb_stage<1>, dim3(b->n_stage_jobs), dim3(64), 0, b->stream, s1);
  }
  if (b->any_eta && from <= 2) {
    if (int rc = enqueue_group_cg(b, 0, L.spec[0])) return rc;
  }
  if (b->any_eta && from <= 3) {
    BatchFinArgs f = fin_args(b, b->rs, b->x, b->st[0], b->S[4], b->S[5], b->S[4], b->S[5]);
    f.ctl = b->ctl;
    BLAUNCH(b, kb_finalize1, dim3(nt), dim3(kTileX), 0, b->stream, f);
  }
  if (from <= 4) {
    BatchStageArgs s2{};
    s2.geom = b->geom;
    s2.rs = b->rs;
    s2.jobs = b->stage_jobs;
    s2.ctl = b->ctl;
    s2.rho_raw = b->rho_raw;
    for (int q = 0; q < 6; ++q) {
      s2.s[q] = b->S[q];
      s2.u[q] = b->W[q];
      s2.o1[q] = b->U[q];
    }
    BLAUNCH(b, kb_stage<2>, dim3(b->n_stage_jobs), dim3(64), 0, b->stream, s2);
  }
  if (b->any_eta && from <= 5) {
    if (int rc = enqueue_group_cg(b, 1, L.spec[1])) return rc;
  }
  if (b->any_eta && from <= 6) {
    BatchFinArgs f = fin_args(b, b->rs, b->x, b->st[1], b->U[4], b->U[5], b->S[4], b->S[5]);
    f.ctl = b->ctl;
    if (fused) BLAUNCH(b, kb_fin2_cfl(b), dim3(nt), dim3(kTileX), 0, b->stream, f, d);
    else BLAUNCH(b, kb_fin2(b), dim3(nt), dim3(kTileX), 0, b->stream, f);
    std::swap(b->U[4], b->S[4]);
    std::swap(b->U[5], b->S[5]);
  }
  return MHD_OK;
}

// The graph of nsteps (even) batch steps for the current B pointers,
// budgets and CFL fusion, captured on first use (least recently used entry
// replaced).
constexpr int kMaxGraphSteps = 64;   // steps per captured graph (graph = 2)

int group_graph(mhd_batch* b, BLoop& L, int nsteps, BatchGraph** out) {
  double* key[4] = {b->U[4], b->U[5], b->S[4], b->S[5]};
  const bool fused = b->fused_cfl != 0;
  BatchGraph* victim = &b->graphs[0];
  for (BatchGraph& g : b->graphs) {
    if (g.exec && std::equal(key, key + 4, g.key) && g.spec[0] == L.spec[0] && g.spec[1] == L.spec[1] &&
        g.fused == fused && g.nsteps == nsteps) {
      g.last_use = ++b->gclock;
      *out = &g;
      return MHD_OK;
    }
    if (!g.exec || (victim->exec && g.last_use < victim->last_use)) victim = &g;
  }
  if (victim->exec) (void)hipGraphExecDestroy(victim->exec);
  victim->exec = nullptr;
  BHIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  int rc = MHD_OK;
  for (int k = 0; k < nsteps && !rc; ++k)       // an even number of swaps: the B pointers end where they started
    rc = enqueue_group_step(b, L, 0);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(b->stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return bfail(b, MHD_EHIP, std::string("batch graph capture: ") + hipGetErrorString(e));
  const hipError_t ei = hipGraphInstantiate(&victim->exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ei != hipSuccess) {
    victim->exec = nullptr;
    return bfail(b, MHD_EHIP, std::string("batch graph instantiate: ") + hipGetErrorString(ei));
  }
  std::copy(key, key + 4, victim->key);
  victim->spec[0] = L.spec[0];
  victim->spec[1] = L.spec[1];
  victim->nsteps = nsteps;
  victim->fused = fused;
  victim->last_use = ++b->gclock;
  *out = victim;
  return MHD_OK;
}

int read_group(mhd_batch* b) {
  BHIP(b, hipMemcpyAsync(b->ctl_host, b->ctl, 4 * sizeof(int), hipMemcpyDeviceToHost, b->stream));
  BHIP(b, hipMemcpyAsync(b->rs_host, b->rs, sizeof(RunState) * b->nruns, hipMemcpyDeviceToHost, b->stream));
  BHIP(b, hipStreamSynchronize(b->stream));
  return MHD_OK;
}

// Longest solve of each slot among the runs' current states.
int read_budgets(mhd_batch* b, int most[2]) {
  BHIP(b, hipMemcpyAsync(b->st_host, b->st[0], sizeof(CgState) * b->nruns, hipMemcpyDeviceToHost, b->stream));
  BHIP(b, hipMemcpyAsync(b->st_host2, b->st[1], sizeof(CgState) * b->nruns, hipMemcpyDeviceToHost, b->stream));
  BHIP(b, hipStreamSynchronize(b->stream));
  most[0] = most[1] = 0;
  for (int r = 0; r < b->nruns; ++r) {
    if (b->st_host[r].done == 1 || b->st_host[r].done == 2) most[0] = std::max(most[0], b->st_host[r].spmv);
    if (b->st_host2[r].done == 1 || b->st_host2[r].done == 2) most[1] = std::max(most[1], b->st_host2[r].spmv);
  }
  return MHD_OK;
}

// A halted solve: continue it with host-checked iterations from the last A
// launched (A(spec + 1)), then record its counts.
int resume_group_solve(mhd_batch* b, BLoop& L, int slot) {
  const BatchCgArgs a = group_cg(b, slot);
  const int nb = b->n_tile_jobs;
  long long maxit = 0;
  for (const RunGeom& g : b->geom_h) maxit = std::max(maxit, g.max_iters);
  long long it = L.spec[slot] + 1;
  for (;;) {
    BLAUNCH(b, kb_pass_b<false>(b), dim3(nb), dim3(kTileX), 0, b->stream, a, (const double*)b->p[it & 1], (int)it);
    ++it;
    BLAUNCH(b, kb_pass_a(b), dim3(nb), dim3(kTileX), 0, b->stream, a, (const double*)b->p[(it - 1) & 1], b->p[it & 1],
            (int)it);
    BHIP(b, hipMemcpyAsync(b->st_host, b->st[slot], sizeof(CgState) * b->nruns, hipMemcpyDeviceToHost, b->stream));
    BHIP(b, hipStreamSynchronize(b->stream));
    bool all = true;
    for (int r = 0; r < b->nruns; ++r) all = all && b->st_host[r].done;
    if (all) break;
    if (it > maxit + 4) return bfail(b, MHD_EHIP, "CG state machine did not terminate");
  }
  int most = 0;
  for (int r = 0; r < b->nruns; ++r)
    if (b->st_host[r].done == 1 || b->st_host[r].done == 2) most = std::max(most, b->st_host[r].spmv);
  L.spec[slot] = std::max(L.spec[slot], most - 1 + b->margin);
  BLAUNCH(b, kb_check_solve, dim3(1), dim3(256), 0, b->stream, (const RunState*)b->rs, (const CgState*)b->st[slot],
          b->nruns, slot, b->o.cg_fail_is_error, b->ctl, b->iters);
  return MHD_OK;
}

// Up to max_batch_steps (< 0: until every run has stopped) steps in groups.
int run_groups(mhd_batch* b, int max_batch_steps, int* steps_done) {
  const int nr = b->nruns;
  BHIP(b, hipMemcpyAsync(b->rs_host, b->rs, sizeof(RunState) * nr, hipMemcpyDeviceToHost, b->stream));
  BHIP(b, hipStreamSynchronize(b->stream));
  std::vector<int> steps0(nr);
  for (int r = 0; r < nr; ++r) steps0[r] = b->rs_host[r].steps;
  BLoop L{};
  for (int k = 0; k < 2; ++k) {
    L.B0[k] = b->U[4 + k];
    L.B1[k] = b->S[4 + k];
    L.spec[k] = std::max(0, b->last_spmv[k] - 1) + b->margin;
  }
  BLAUNCH(b, kb_ctl_reset, dim3(1), dim3(64), 0, b->stream, b->ctl);
  const bool fused = b->any_eta && b->fused_cfl;
  if (fused && !b->cfl_fresh) {
    const BatchDiagArgs d = group_diag(b);
    BLAUNCH(b, kb_cfl, dim3(b->n_tile_jobs), dim3(kTileX), 0, b->stream, d);
  }
  b->cfl_fresh = false;
  long long done = 0;
  int rc = MHD_OK;
  for (;;) {
    if (max_batch_steps >= 0 && done >= max_batch_steps) break;
    long long G = b->group;
    if (max_batch_steps >= 0) G = std::min<long long>(G, max_batch_steps - done);
    if (b->graph != 2) {
      // no more steps than the longest remaining step limit when every active
      // run has one (later steps would be launches that return at once); a
      // whole-group graph keeps its size, so that its capture is reused
      long long most = 0;
      bool limited = true;
      for (int r = 0; r < nr && limited; ++r) {
        if (!b->rs_host[r].active) continue;
        const int ms = b->geom_h[r].max_steps;
        limited = ms >= 0;
        most = std::max<long long>(most, (long long)ms - b->rs_host[r].steps);
      }
      if (limited) G = std::min(G, std::max(most, 1LL));
    }
    long long k = 0;
    if (b->graph == 2) {                   // the group's even part as one graph launch (chunks of at most 64 steps)
      while (k + 2 <= G) {
        const int n = (int)std::min<long long>((G - k) & ~1LL, kMaxGraphSteps);
        BatchGraph* g = nullptr;
        if ((rc = group_graph(b, L, n, &g))) return rc;
        BHIP(b, hipGraphLaunch(g->exec, b->stream));
        k += n;
      }
    } else if (b->graph == 1) {
      for (; k + 2 <= G; k += 2) {
        BatchGraph* g = nullptr;
        if ((rc = group_graph(b, L, 2, &g))) return rc;
        BHIP(b, hipGraphLaunch(g->exec, b->stream));
      }
    }
    for (; k < G; ++k)
      if ((rc = enqueue_group_step(b, L, 0))) return rc;
    if ((rc = read_group(b))) return rc;
    long long began = b->ctl_host[kCtlSteps];
    while (b->ctl_host[kCtlHalt] == kHaltSolve1 || b->ctl_host[kCtlHalt] == kHaltSolve2) {
      const int slot = b->ctl_host[kCtlHalt] == kHaltSolve1 ? 0 : 1;
      const long long kh = began - 1;                       // steps of the group finished before the halted one
      set_bparity(b, L, done + kh);
      BLAUNCH(b, kb_ctl_reset, dim3(1), dim3(64), 0, b->stream, b->ctl);
      if ((rc = resume_group_solve(b, L, slot))) return rc;
      if ((rc = enqueue_group_step(b, L, slot == 0 ? 3 : 6))) return rc;
      for (long long q = kh + 1; q < G; ++q)
        if ((rc = enqueue_group_step(b, L, 0))) return rc;
      if ((rc = read_group(b))) return rc;
      began = kh + 1 + b->ctl_host[kCtlSteps];
    }
    const int h = b->ctl_host[kCtlHalt];
    if (h == kHaltNonfinite) return bfail(b, MHD_ENONFINITE, "non-finite wave speed (batched group of steps)");
    if (h == kHaltCgLimit1 || h == kHaltCgLimit2)
      return bfail(b, MHD_ECG, "a run's CG solve reached the iteration limit (batched group of steps)");
    done += G;
    set_bparity(b, L, done);
    int act = 0;
    for (int r = 0; r < nr; ++r) act += b->rs_host[r].active;
    if (b->any_eta) {
      int most[2];
      if ((rc = read_budgets(b, most))) return rc;
      for (int q = 0; q < 2; ++q)
        if (most[q] > 0) {
          L.spec[q] = most[q] - 1 + b->margin;
          b->last_spmv[q] = most[q];
        }
    }
    BLAUNCH(b, kb_ctl_reset, dim3(1), dim3(64), 0, b->stream, b->ctl);
    if (act == 0) break;
  }
  BHIP(b, hipMemcpyAsync(b->iters_host, b->iters, 2 * sizeof(int) * nr, hipMemcpyDeviceToHost, b->stream));
  BHIP(b, hipStreamSynchronize(b->stream));
  int stepped = 0;
  for (int r = 0; r < nr; ++r) {
    const int ds = b->rs_host[r].steps - steps0[r];
    stepped = std::max(stepped, ds);
    if (ds > 0) {
      b->cg_iters_last[2 * r] = b->iters_host[2 * r];
      b->cg_iters_last[2 * r + 1] = b->iters_host[2 * r + 1];
    }
  }
  b->steps_taken += stepped;
  b->cfl_fresh = fused;
  if (steps_done) *steps_done = stepped;
  return MHD_OK;
}

}  // namespace

extern "C" {

int mhd_batch_create(int nruns, const int* nx, const int* ny, const double* dx, const double* dy, const double* eta,
                     const mhd_opts* opts, mhd_batch** out) {
  if (!out) return bfail(nullptr, MHD_EINVAL, "out is NULL");
  *out = nullptr;
  if (nruns < 1 || !nx || !ny || !dx || !dy || !eta) return bfail(nullptr, MHD_EINVAL, "bad run arrays");
  mhd_opts o;
  mhd_default_opts(&o);
  if (opts) o = *opts;
  if (o.gamma != 5.0 / 3.0) return bfail(nullptr, MHD_EINVAL, "only gamma = 5/3 is compiled in");
  if (!(o.cfl > 0) || !(o.cg_tol >= 0) || o.march_rows < 0) return bfail(nullptr, MHD_EINVAL, "bad options");
  const bool bc_ok = [](int v) { return v == MHD_BC_PERIODIC || v == MHD_BC_OUTFLOW; }(o.bc_x) &&
                     [](int v) { return v == MHD_BC_PERIODIC || v == MHD_BC_OUTFLOW; }(o.bc_y);
  if (!bc_ok) return bfail(nullptr, MHD_EINVAL, "bc_x and bc_y must be MHD_BC_PERIODIC or MHD_BC_OUTFLOW");
  long long total = 0;
  for (int r = 0; r < nruns; ++r) {
    if (nx[r] < 4 || ny[r] < 4) return bfail(nullptr, MHD_EINVAL, "run " + std::to_string(r) + ": nx, ny must be >= 4");
    if (!(dx[r] > 0) || !(dy[r] > 0) || !std::isfinite(dx[r]) || !std::isfinite(dy[r]))
      return bfail(nullptr, MHD_EINVAL, "run " + std::to_string(r) + ": bad dx or dy");
    if (!(eta[r] >= 0) || !std::isfinite(eta[r])) return bfail(nullptr, MHD_EINVAL, "run " + std::to_string(r) + ": bad eta");
    total += (long long)nx[r] * ny[r];
  }
  if (total > (1LL << 31)) return bfail(nullptr, MHD_EINVAL, "more than 2^31 cells in the batch");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return bfail(nullptr, MHD_EHIP, "no HIP device available");

  mhd_batch* b = new mhd_batch();
  b->nruns = nruns;
  b->o = o;
  b->geom_h.resize(nruns);
  b->cg_iters_last.assign(2 * nruns, 0);
  for (int r = 0; r < nruns; ++r) {
    RunGeom& g = b->geom_h[r];
    g = RunGeom{};
    g.nx = nx[r];
    g.ny = ny[r];
    g.dx = dx[r];
    g.dy = dy[r];
    g.inv_dx = 1.0 / dx[r];
    g.inv_dy = 1.0 / dy[r];
    g.eta = eta[r];
    g.cfl = o.cfl;
    g.t_end = INFINITY;
    g.dt_fixed = 0.0;
    g.max_steps = -1;
    g.max_iters = (o.cg_max_iter > 0) ? o.cg_max_iter : 2LL * nx[r] * ny[r];
    g.open = (o.bc_x == MHD_BC_OUTFLOW ? 1 : 0) | (o.bc_y == MHD_BC_OUTFLOW ? 2 : 0);
    b->any_eta = b->any_eta || eta[r] > 0.0;
  }
  std::vector<int2> sj, tj;
  // capacity: the smallest candidate shapes give the most jobs and tiles
  b->stage_rows = kStageRowChoices[0];
  b->tile_rows = kTileRowChoices[0];
  plan(b, sj, tj);
  b->cap_stage_jobs = std::max(sj.size(), (size_t)1);
  b->cap_tiles = std::max(tj.size(), (size_t)1);
  b->stage_rows = (o.march_rows > 0) ? o.march_rows : 32;
  b->tile_rows = 32;
  plan(b, sj, tj);
  b->cap_stage_jobs = std::max(b->cap_stage_jobs, sj.size());
  b->cap_tiles = std::max(b->cap_tiles, tj.size());
  auto bail = [&](const std::string& m, int code) {
    g_batch_create_error = m;
    mhd_batch_destroy(b);
    return code;
  };
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) return bail("hipStreamCreate failed", MHD_EHIP);
  const size_t fb = (size_t)b->total * sizeof(double);
  for (int q = 0; q < 6; ++q)
    if (hipMalloc(&b->U[q], fb) != hipSuccess || hipMalloc(&b->S[q], fb) != hipSuccess || hipMalloc(&b->W[q], fb) != hipSuccess)
      return bail("device allocation failed", MHD_ENOMEM);
  const size_t pb = b->cap_tiles * sizeof(double);
  if (hipMalloc(&b->x, fb) != hipSuccess || hipMalloc(&b->r, fb) != hipSuccess || hipMalloc(&b->p[0], fb) != hipSuccess ||
      hipMalloc(&b->rho_raw, fb) != hipSuccess ||
      hipMalloc(&b->p[1], fb) != hipSuccess || hipMalloc(&b->part, 2 * pb) != hipSuccess ||
      hipMalloc(&b->part_s, 3 * pb) != hipSuccess || hipMalloc(&b->part_a, pb) != hipSuccess ||
      hipMalloc(&b->part_b, 2 * pb) != hipSuccess || hipMalloc(&b->st[0], sizeof(CgState) * nruns) != hipSuccess ||
      hipMalloc(&b->st[1], sizeof(CgState) * nruns) != hipSuccess || hipMalloc(&b->geom, sizeof(RunGeom) * nruns) != hipSuccess ||
      hipMalloc(&b->rs, sizeof(RunState) * nruns) != hipSuccess ||
      hipMalloc(&b->stage_jobs, sizeof(int2) * b->cap_stage_jobs) != hipSuccess ||
      hipMalloc(&b->tile_jobs, sizeof(int2) * b->cap_tiles) != hipSuccess)
    return bail("device allocation failed", MHD_ENOMEM);
  long long largest = 0;
  for (int r = 0; r < nruns; ++r) largest = std::max(largest, (long long)nx[r] * ny[r]);
  if (hipHostMalloc(&b->rs_host, sizeof(RunState) * nruns) != hipSuccess ||
      hipHostMalloc(&b->st_host, sizeof(CgState) * nruns) != hipSuccess ||
      hipHostMalloc(&b->st_host2, sizeof(CgState) * nruns) != hipSuccess ||
      hipHostMalloc(&b->ctl_host, 4 * sizeof(int)) != hipSuccess ||
      hipHostMalloc(&b->iters_host, 2 * sizeof(int) * nruns) != hipSuccess ||
      hipHostMalloc(&b->stage, sizeof(double) * largest) != hipSuccess)
    return bail("pinned allocation failed", MHD_ENOMEM);
  if (hipMalloc(&b->ctl, 4 * sizeof(int)) != hipSuccess || hipMalloc(&b->iters, 2 * sizeof(int) * nruns) != hipSuccess)
    return bail("device allocation failed", MHD_ENOMEM);
  if (hipMemsetAsync(b->ctl, 0, 4 * sizeof(int), b->stream) != hipSuccess ||
      hipMemsetAsync(b->iters, 0, 2 * sizeof(int) * nruns, b->stream) != hipSuccess)
    return bail("hipMemset failed", MHD_EHIP);
  b->n_stage_jobs = (int)sj.size();
  b->n_tile_jobs = (int)tj.size();
  for (int q = 0; q < 6; ++q)
    if (hipMemsetAsync(b->U[q], 0, fb, b->stream) != hipSuccess) return bail("hipMemset failed", MHD_EHIP);
  if (hipMemsetAsync(b->rho_raw, 0, fb, b->stream) != hipSuccess) return bail("hipMemset failed", MHD_EHIP);
  std::memset(b->rs_host, 0, sizeof(RunState) * nruns);
  if (hipMemcpyAsync(b->geom, b->geom_h.data(), sizeof(RunGeom) * nruns, hipMemcpyHostToDevice, b->stream) != hipSuccess ||
      hipMemcpyAsync(b->rs, b->rs_host, sizeof(RunState) * nruns, hipMemcpyHostToDevice, b->stream) != hipSuccess ||
      hipMemcpyAsync(b->stage_jobs, sj.data(), sizeof(int2) * sj.size(), hipMemcpyHostToDevice, b->stream) != hipSuccess ||
      hipMemcpyAsync(b->tile_jobs, tj.data(), sizeof(int2) * tj.size(), hipMemcpyHostToDevice, b->stream) != hipSuccess ||
      hipStreamSynchronize(b->stream) != hipSuccess)
    return bail("device init failed", MHD_EHIP);
  *out = b;
  return MHD_OK;
}

int mhd_batch_destroy(mhd_batch* b) {
  if (!b) return MHD_EINVAL;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  destroy_batch_adjoint(b);
  for (int q = 0; q < 6; ++q) {
    if (b->U[q]) (void)hipFree(b->U[q]);
    if (b->S[q]) (void)hipFree(b->S[q]);
    if (b->W[q]) (void)hipFree(b->W[q]);
  }
  drop_graphs(b);
  void* bufs[] = {b->x, b->r, b->p[0], b->p[1], b->part, b->part_s, b->part_a, b->part_b, b->st[0], b->st[1],
                  b->geom, b->rs, b->stage_jobs, b->tile_jobs, b->ctl, b->iters, b->rho_raw};
  for (void* p : bufs)
    if (p) (void)hipFree(p);
  if (b->rs_host) (void)hipHostFree(b->rs_host);
  if (b->st_host) (void)hipHostFree(b->st_host);
  if (b->st_host2) (void)hipHostFree(b->st_host2);
  if (b->ctl_host) (void)hipHostFree(b->ctl_host);
  if (b->iters_host) (void)hipHostFree(b->iters_host);
  if (b->stage) (void)hipHostFree(b->stage);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return MHD_OK;
}

const char* mhd_batch_last_error(const mhd_batch* b) { return b ? b->err.c_str() : g_batch_create_error.c_str(); }

int mhd_batch_set_schedule(mhd_batch* b, const double* t_end, const int* max_steps, const double* dt_fixed) {
  if (b) b->cfl_fresh = false;
  if (!b) return MHD_EINVAL;
  for (int r = 0; r < b->nruns; ++r) {
    RunGeom& g = b->geom_h[r];
    if (t_end) {
      if (std::isnan(t_end[r])) return bfail(b, MHD_EINVAL, "t_end is NaN");
      g.t_end = t_end[r];
    }
    if (max_steps) g.max_steps = max_steps[r];
    if (dt_fixed) {
      if (!(dt_fixed[r] >= 0) || !std::isfinite(dt_fixed[r])) return bfail(b, MHD_EINVAL, "bad dt_fixed");
      g.dt_fixed = dt_fixed[r];
    }
  }
  BHIP(b, hipMemcpyAsync(b->geom, b->geom_h.data(), sizeof(RunGeom) * b->nruns, hipMemcpyHostToDevice, b->stream));
  BHIP(b, hipStreamSynchronize(b->stream));
  return MHD_OK;
}

int mhd_batch_upload(mhd_batch* b, int run, const double* rho, const double* mx, const double* my, const double* E,
                     const double* bxf, const double* byf) {
  int rc = check_run(b, run);
  if (rc) return rc;
  b->cfl_fresh = false;
  const double* src[6] = {rho, mx, my, E, bxf, byf};
  for (int q = 0; q < 6; ++q)
    if (!src[q]) return bfail(b, MHD_EINVAL, "mhd_batch_upload: NULL field");
  const RunGeom& g = b->geom_h[run];
  for (int q = 0; q < 6; ++q)
    if ((rc = stage_in(b, run, src[q], b->U[q]))) return rc;
  const long long n = (long long)g.nx * g.ny;
  const int blocks = (int)std::min<long long>((n + 255) / 256, 4096);
  BLAUNCH(b, k_floor_rho, dim3(blocks), dim3(256), 0, b->stream, b->U[0] + g.base, (int)n);
  // the run's first CFL step reads the primitives of the floored state
  BLAUNCH(b, k_raw_reset, dim3(blocks), dim3(256), 0, b->stream, (const double*)b->U[0] + g.base, b->rho_raw + g.base,
          (int)n);
  // a fresh state restarts the run's driver
  RunState z{};
  BHIP(b, hipMemcpyAsync(b->rs + run, &z, sizeof(RunState), hipMemcpyHostToDevice, b->stream));
  BHIP(b, hipStreamSynchronize(b->stream));
  return MHD_OK;
}

int mhd_batch_download(mhd_batch* b, int run, double* rho, double* mx, double* my, double* E, double* bxf, double* byf) {
  int rc = check_run(b, run);
  if (rc) return rc;
  double* dst[6] = {rho, mx, my, E, bxf, byf};
  for (int q = 0; q < 6; ++q)
    if (dst[q] && (rc = stage_out(b, run, b->U[q], dst[q]))) return rc;
  return MHD_OK;
}

int mhd_batch_step(mhd_batch* b, int* n_active) {
  if (!b) return MHD_EINVAL;
  const BatchTableau& T = b->tab;
  const int nt = b->n_tile_jobs;
  // driver: CFL partials, then dt / operator / activity per run
  BatchDiagArgs d{};
  d.geom = b->geom;
  d.rs = b->rs;
  d.jobs = b->tile_jobs;
  for (int q = 0; q < 6; ++q) d.s[q] = b->U[q];
  d.part = b->part;
  d.rho_raw = b->rho_raw;
  // the previous step's final recombination formed the partials already
  const bool fresh = b->cfl_fresh;
  b->cfl_fresh = false;
  if (!fresh) BLAUNCH(b, kb_cfl, dim3(nt), dim3(kTileX), 0, b->stream, d);
  BLAUNCH(b, kb_dt, dim3(b->nruns), dim3(256), 0, b->stream, (const RunGeom*)b->geom, b->rs, (const double*)b->part,
          b->st[0], b->st[1], b->nruns, step_factors(T), 0, nullptr);
  BHIP(b, hipMemcpyAsync(b->rs_host, b->rs, sizeof(RunState) * b->nruns, hipMemcpyDeviceToHost, b->stream));
  // stage 1: W = U + dt delta f1 (into W), U2* = U + dt gamma f1 (into S)
  BatchStageArgs s1{};
  s1.geom = b->geom;
  s1.rs = b->rs;
  s1.jobs = b->stage_jobs;
  for (int q = 0; q < 6; ++q) {
    s1.s[q] = b->U[q];
    s1.o1[q] = b->W[q];
    s1.o2[q] = b->S[q];
  }
  BLAUNCH(b, kb_stage<1>, dim3(b->n_stage_jobs), dim3(64), 0, b->stream, s1);
  if (b->any_eta) {
    CgCall cc;
    cc.slot = 0;
    cc.bx = b->S[4];
    cc.by = b->S[5];
    cc.x = b->x;
    cc.st = b->st[0];
    cc.rs = b->rs;
    int rc = cg_batch(b, cc);
    if (rc) return rc;
    BatchFinArgs f = fin_args(b, b->rs, b->x, b->st[0], b->S[4], b->S[5], b->S[4], b->S[5]);
    BLAUNCH(b, kb_finalize1, dim3(nt), dim3(kTileX), 0, b->stream, f);
  }
  // stage 2: U3* = W + dt (1 - delta) L(S) [+ dt (1 - gamma) g2] into U
  BatchStageArgs s2{};
  s2.geom = b->geom;
  s2.rs = b->rs;
  s2.jobs = b->stage_jobs;
  s2.rho_raw = b->rho_raw;
  for (int q = 0; q < 6; ++q) {
    s2.s[q] = b->S[q];
    s2.u[q] = b->W[q];
    s2.o1[q] = b->U[q];
  }
  BLAUNCH(b, kb_stage<2>, dim3(b->n_stage_jobs), dim3(64), 0, b->stream, s2);
  if (b->any_eta) {
    CgCall cc;
    cc.slot = 1;
    cc.bx = b->U[4];
    cc.by = b->U[5];
    cc.x = b->x;
    cc.st = b->st[1];
    cc.rs = b->rs;
    int rc = cg_batch(b, cc);
    if (rc) return rc;
    // B^{n+1} = B3* + dt gamma g3(B3) into S.B, then S.B becomes the state's B.
    // Runs with eta = 0 have no implicit step: their B is copied by the
    // finalizer as well (x = 0 there, so g3 = 0 and c g3 = 0).
    BatchFinArgs f = fin_args(b, b->rs, b->x, b->st[1], b->U[4], b->U[5], b->S[4], b->S[5]);
    BLAUNCH(b, kb_fin2_cfl(b), dim3(nt), dim3(kTileX), 0, b->stream, f, d);
    std::swap(b->U[4], b->S[4]);
    std::swap(b->U[5], b->S[5]);
  }
  BHIP(b, hipStreamSynchronize(b->stream));
  int act = 0;
  for (int r = 0; r < b->nruns; ++r) {
    if (b->rs_host[r].nonfinite)
      return bfail(b, MHD_ENONFINITE, "run " + std::to_string(r) + ": non-finite wave speed");
    act += b->rs_host[r].active;
  }
  b->steps_taken++;
  b->cfl_fresh = b->any_eta;
  if (n_active) *n_active = act;
  return MHD_OK;
}

int mhd_batch_tune(mhd_batch* b, int group_steps, int spec_margin, int graph, int fused_cfl) {
  if (!b) return MHD_EINVAL;
  if (group_steps < 0 || group_steps > 4096 || spec_margin < 0 || graph < 0 || graph > 2 ||
      (fused_cfl != 0 && fused_cfl != 1))
    return bfail(b, MHD_EINVAL, "mhd_batch_tune: group_steps in [0, 4096], spec_margin >= 0, graph 0, 1 or 2, "
                                "fused_cfl 0 or 1");
  b->group = group_steps;
  b->margin = spec_margin;
  b->graph = graph;
  if (b->fused_cfl != fused_cfl) b->cfl_fresh = false;
  b->fused_cfl = fused_cfl;
  return MHD_OK;
}

int mhd_batch_get_tuning(mhd_batch* b, int* stage_rows, int* tile_rows, int* group_steps, int* graph,
                         int* fused_cfl) {
  if (!b) return MHD_EINVAL;
  if (stage_rows) *stage_rows = b->stage_rows;
  if (tile_rows) *tile_rows = b->tile_rows;
  if (group_steps) *group_steps = b->group;
  if (graph) *graph = b->graph;
  if (fused_cfl) *fused_cfl = b->fused_cfl;
  return MHD_OK;
}

int mhd_batch_run(mhd_batch* b, int max_batch_steps, int* steps_done) {
  if (!b) return MHD_EINVAL;
  if (b->group > 0) return run_groups(b, max_batch_steps, steps_done);
  int k = 0;
  for (; max_batch_steps < 0 || k < max_batch_steps; ++k) {
    int act = 0;
    int rc = mhd_batch_step(b, &act);
    if (rc) return rc;
    if (act == 0) break;
  }
  if (steps_done) *steps_done = k;
  return MHD_OK;
}

int mhd_batch_autotune(mhd_batch* b, int trial_steps, int* stage_rows, int* tile_rows) {
  if (!b) return MHD_EINVAL;
  if (trial_steps < 1) return bfail(b, MHD_EINVAL, "trial_steps must be >= 1");
  // save the state, the driver state and the host-side step bookkeeping
  const size_t fb = (size_t)b->total * sizeof(double);
  double* saved[6] = {};
  RunState* rs_saved = nullptr;
  auto release = [&]() {
    for (double* p : saved)
      if (p) (void)hipFree(p);
    if (rs_saved) (void)hipFree(rs_saved);
  };
  for (int q = 0; q < 6; ++q)
    if (hipMalloc(&saved[q], fb) != hipSuccess) {
      release();
      return bfail(b, MHD_ENOMEM, "autotune: device allocation failed");
    }
  if (hipMalloc(&rs_saved, sizeof(RunState) * b->nruns) != hipSuccess) {
    release();
    return bfail(b, MHD_ENOMEM, "autotune: device allocation failed");
  }
  for (int q = 0; q < 6; ++q)
    (void)hipMemcpyAsync(saved[q], b->U[q], fb, hipMemcpyDeviceToDevice, b->stream);
  (void)hipMemcpyAsync(rs_saved, b->rs, sizeof(RunState) * b->nruns, hipMemcpyDeviceToDevice, b->stream);
  const int spmv_saved[3] = {b->last_spmv[0], b->last_spmv[1], b->last_spmv[2]};
  int spmv_seen[3] = {spmv_saved[0], spmv_saved[1], spmv_saved[2]};
  const std::vector<int> its_saved = b->cg_iters_last;
  const long long steps_saved = b->steps_taken;
  std::vector<RunGeom> geom_saved = b->geom_h;
  // the trial runs every run for trial_steps steps regardless of its schedule
  for (RunGeom& g : b->geom_h) {
    g.t_end = INFINITY;
    g.max_steps = -1;
  }
  double best = INFINITY;
  int best_s = b->stage_rows, best_t = b->tile_rows;
  const int graph_saved = b->graph, fused_saved = b->fused_cfl;
  int best_g = b->graph, best_f = b->fused_cfl;
  int rc = MHD_OK;
  // time trial_steps steps of the current configuration after a warm-up
  // (two steps: a graph configuration captures its pair there), then put the
  // state back
  auto trial = [&](double& total) -> int {
    total = 0.0;
    int r = MHD_OK;
    if (b->group > 0) {
      int nd = 0;
      if ((r = mhd_batch_run(b, 2, &nd))) return r;
      if (hipStreamSynchronize(b->stream) != hipSuccess) return bfail(b, MHD_EHIP, "autotune: sync failed");
      auto t0 = std::chrono::steady_clock::now();
      r = mhd_batch_run(b, trial_steps, &nd);
      auto t1 = std::chrono::steady_clock::now();
      total = std::chrono::duration<double>(t1 - t0).count();
    } else {
      for (int k = 0; k < trial_steps + 1 && r == MHD_OK; ++k) {
        if (hipStreamSynchronize(b->stream) != hipSuccess) return bfail(b, MHD_EHIP, "autotune: sync failed");
        auto t0 = std::chrono::steady_clock::now();
        r = mhd_batch_step(b, nullptr);
        auto t1 = std::chrono::steady_clock::now();
        if (k > 0) total += std::chrono::duration<double>(t1 - t0).count();   // the first step warms the shape up
      }
    }
    for (int q = 0; q < 6; ++q)
      (void)hipMemcpyAsync(b->U[q], saved[q], fb, hipMemcpyDeviceToDevice, b->stream);
    (void)hipMemcpyAsync(b->rs, rs_saved, sizeof(RunState) * b->nruns, hipMemcpyDeviceToDevice, b->stream);
    for (int k = 0; k < 3; ++k) {
      spmv_seen[k] = std::max(spmv_seen[k], b->last_spmv[k]);
      b->last_spmv[k] = spmv_saved[k];
    }
    b->cfl_fresh = false;
    return r;
  };
  // launch shapes (in the configured launch mode)
  for (int sr : kStageRowChoices) {
    for (int tr : kTileRowChoices) {
      rc = replan(b, sr, tr);
      if (rc) break;
      double total = 0.0;
      rc = trial(total);
      if (rc) break;
      if (total < best) {
        best = total;
        best_s = sr;
        best_t = tr;
      }
    }
    if (rc) break;
  }
  // stage fusion and graph replay of the device-resident groups, with the
  // best shape: the next step's CFL partials formed by the final
  // recombination or by a separate pass; direct launches, captured pairs or
  // each group as one captured graph
  if (rc == MHD_OK && b->group > 0 && b->any_eta) {
    rc = replan(b, best_s, best_t);
    for (int f = 1; f >= 0 && rc == MHD_OK; --f)
      for (int g = 0; g <= 2 && rc == MHD_OK; ++g) {
        b->fused_cfl = f;
        b->graph = g;
        double total = 0.0;
        rc = trial(total);
        if (rc) break;
        if (total < best) {
          best = total;
          best_f = f;
          best_g = g;
        }
      }
  }
  b->graph = (rc == MHD_OK) ? best_g : graph_saved;
  b->fused_cfl = (rc == MHD_OK) ? best_f : fused_saved;
  b->geom_h = geom_saved;
  // the trials' longest solves seed the device-resident groups' CG budgets (a
  // budget only sets how many iterations a group launches before the device
  // checks; one too short halts the group for a host-checked resume)
  for (int k = 0; k < 3; ++k) b->last_spmv[k] = spmv_seen[k];
  b->cg_iters_last = its_saved;
  b->steps_taken = steps_saved;
  int rc2 = replan(b, best_s, best_t);
  release();
  if (rc) return rc;
  if (rc2) return rc2;
  if (stage_rows) *stage_rows = best_s;
  if (tile_rows) *tile_rows = best_t;
  return MHD_OK;
}

int mhd_batch_run_info(mhd_batch* b, int run, int* steps, double* t, double* last_dt, int* cg_iters) {
  int rc = check_run(b, run);
  if (rc) return rc;
  RunState s;
  BHIP(b, hipMemcpyAsync(&s, b->rs + run, sizeof(RunState), hipMemcpyDeviceToHost, b->stream));
  BHIP(b, hipStreamSynchronize(b->stream));
  if (steps) *steps = s.steps;
  if (t) *t = s.t;
  if (last_dt) *last_dt = s.dt;
  if (cg_iters) {
    cg_iters[0] = b->cg_iters_last[2 * run];
    cg_iters[1] = b->cg_iters_last[2 * run + 1];
  }
  return MHD_OK;
}

int mhd_batch_synchronize(mhd_batch* b) {
  if (!b) return MHD_EINVAL;
  BHIP(b, hipStreamSynchronize(b->stream));
  return MHD_OK;
}

}  // extern "C"
