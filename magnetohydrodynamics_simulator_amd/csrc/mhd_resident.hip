// mhd_resident.hip -- the device-resident driver loop of a single context
// (C ABI mhd_run_resident, mhd_resident_tune).
//
// Same semantics as mhd_run (the loop of src/integrator_main.cpp:52-112 of
// the reference: CFL dt or the constant override, the clamp of the last step
// to t_end, the operator rebuild when dt moves by more than 1e-10 relative,
// IMEX), with the per-step decisions taken on the device:
//   kb_dt (one run)  dt, clamp, rebuild rule, t, step count, activity
//   k_stage_res<1>, CG 1, k_finalize1_res, k_stage_res<2>, CG 2,
//   k_finalize2_res (final recombination + step log + the CFL partials of
//   the new state, fused: the separate CFL pass of mhd_run re-reads the
//   whole state); eta = 0: k_cfl_res (step log + CFL partials)
// Every kernel reads its coefficients from the device driver state, so the
// host enqueues a group of steps and reads the driver state back once per
// group instead of three times per step (the CFL value and the two CG
// convergence checks of mhd_run).
//
// CG: each solve gets a budget of launched iterations -- the most products
// of the slot's solves in the previous group, plus a margin (default 0: a
// smooth run's counts change rarely, and each launched pass after convergence
// still costs a dispatch); the passes of a solve that has converged return at
// once.  A finaliser that finds its solve
// unfinished halts the loop (RunState.halt): every later kernel returns at
// once, the host finds the halt at the group's read-back, continues that
// solve with host-checked iterations, finishes the step and re-enters the
// loop.  The arithmetic is that of mhd_run kernel for kernel, so the states,
// time steps and iteration counts are identical.
//
// Snapshot cadence: with snapshot_every > 0 groups end on multiples of it and
// the caller's callback runs there (frame 0 before the first step, as the
// reference writes its initial snapshot), with the stream idle, so it may
// download or write the state (mhd_write_snapshot).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>

#include "../../include/mhd_mi355x.h"
#include "mhd_context.hpp"
#include "mhd_kernels.hpp"

using namespace mhd;

namespace mhd_host {

constexpr int kLogCap = 256;   // step-log ring entries (>= the largest group)
constexpr int kGraphCache = 4; // captured two-step graphs kept per context

// Two steps captured as a graph: valid for the B pointers, CG budgets and
// dt mode it was captured with (every other argument is fixed for the
// context's life).
struct StepGraph {
  hipGraphExec_t exec = nullptr;
  double* key[4] = {};         // U[4], U[5], S[4], S[5] at capture
  int spec[2] = {0, 0};
  int nsteps = 0;              // steps captured (even)
  bool fixed_dt = false;
  unsigned long long last_use = 0;
};

struct Resident {
  RunGeom* geom = nullptr;     // device, one run
  RunState* rs = nullptr;      // device driver state
  CgState* st = nullptr;       // device, the step's two solves
  StepLog* log = nullptr;      // device ring of kLogCap entries
  RunState* rs_host = nullptr;     // pinned
  CgState* st_host = nullptr;      // pinned, 2
  StepLog* log_host = nullptr;     // pinned, kLogCap
  int group = 16;              // steps enqueued between two read-backs
  int margin = 0;              // extra launched iterations per solve
  int fixed = -1;              // >= 0: fixed launched iterations per solve
  int graph = 0;               // 1: replay pairs of steps as captured graphs
  StepGraph graphs[kGraphCache];
  unsigned long long clock = 0;
};

void destroy_resident(mhd_ctx* c) {
  Resident* R = c->res;
  if (!R) return;
  if (R->geom) (void)hipFree(R->geom);
  if (R->rs) (void)hipFree(R->rs);
  if (R->st) (void)hipFree(R->st);
  if (R->log) (void)hipFree(R->log);
  if (R->rs_host) (void)hipHostFree(R->rs_host);
  if (R->st_host) (void)hipHostFree(R->st_host);
  if (R->log_host) (void)hipHostFree(R->log_host);
  for (StepGraph& g : R->graphs)
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
  delete R;
  c->res = nullptr;
}

}  // namespace mhd_host

using namespace mhd_host;

namespace {

int ensure_resident(mhd_ctx* c) {
  if (c->res) return MHD_OK;
  Resident* R = new Resident();
  c->res = R;
  if (hipMalloc(&R->geom, sizeof(RunGeom)) != hipSuccess || hipMalloc(&R->rs, sizeof(RunState)) != hipSuccess ||
      hipMalloc(&R->st, 2 * sizeof(CgState)) != hipSuccess ||
      hipMalloc(&R->log, kLogCap * sizeof(StepLog)) != hipSuccess ||
      hipHostMalloc(&R->rs_host, sizeof(RunState)) != hipSuccess ||
      hipHostMalloc(&R->st_host, 2 * sizeof(CgState)) != hipSuccess ||
      hipHostMalloc(&R->log_host, kLogCap * sizeof(StepLog)) != hipSuccess) {
    destroy_resident(c);
    return fail(c, MHD_ENOMEM, "mhd_run_resident: device allocation failed");
  }
  return MHD_OK;
}

// Loop-wide constants of the step's launches.
struct Loop {
  Resident* R;
  double* B0[2];      // the state's face B after `base` steps of the loop (U[4], U[5])
  double* B1[2];      // the other pair (S[4], S[5])
  long long base;     // steps at the last rebase (the start, or the last snapshot callback)
  int spec[2];        // launched iterations per solve (A(1) B(1) ... A(spec) B(spec) A(spec + 1))
  bool fixed_dt;
};

// The context's B pointers after `completed` steps of this loop (each step
// with a resistive solve ends with the swap of the final recombination).
void set_parity(mhd_ctx* c, const Loop& L, long long completed) {
  const bool odd = (c->eta > 0.0) && ((completed - L.base) & 1);
  for (int k = 0; k < 2; ++k) {
    c->U[4 + k] = odd ? L.B1[k] : L.B0[k];
    c->S[4 + k] = odd ? L.B0[k] : L.B1[k];
  }
}

CgArgs res_cg(mhd_ctx* c, const Loop& L, const double* bx, const double* by, int slot) {
  CgArgs a = cg_args(c, bx, by);
  a.st = L.R->st + slot;
  return a;
}

// Launched iterations of a solve: start, B(0), then A(1) B(1) ... A(spec) B(spec) A(spec + 1).
int enqueue_cg(mhd_ctx* c, const CgArgs& a, const RunState* rs, int spec) {
  const int nb = c->tiles;
  const bool open = a.open_x | a.open_y;
  auto k_start = open ? k_cg_start_res<true> : k_cg_start_res<false>;
  auto k_b0 = open ? k_cg_pass_b_res<true, true> : k_cg_pass_b_res<true, false>;
  auto k_a = open ? k_cg_pass_a_res<true> : k_cg_pass_a_res<false>;
  auto k_b = open ? k_cg_pass_b_res<false, true> : k_cg_pass_b_res<false, false>;
  int rc = launch_k(c, 2, k_start, dim3(nb), dim3(kTileX), a, rs);
  if (rc) return rc;
  rc = launch_k(c, 4, k_b0, dim3(nb), dim3(kTileX), a, rs, (const double*)c->p[0], 0);
  if (rc) return rc;
  for (int it = 1; it <= spec + 1; ++it) {
    rc = launch_k(c, 3, k_a, dim3(nb), dim3(kTileX), a, rs, (const double*)c->p[(it - 1) & 1], c->p[it & 1], it);
    if (rc) return rc;
    if (it == spec + 1) break;
    rc = launch_k(c, 4, k_b, dim3(nb), dim3(kTileX), a, rs, (const double*)c->p[it & 1], it);
    if (rc) return rc;
  }
  return MHD_OK;
}

// Phases of a step: 0 kb_dt, 1 stage 1, 2 CG 1, 3 finaliser 1, 4 stage 2,
// 5 CG 2, 6 finaliser 2 (+ the B swap; it also writes the step log and the
// CFL partials); without a resistive solve, 7 step log and CFL partials.
int enqueue_step(mhd_ctx* c, Loop& L, int from) {
  Resident* R = L.R;
  const Tableau& T = c->tab;
  const bool res = c->eta > 0.0;
  int rc;
  if (from <= 0) {
    rc = launch_k(c, 8, kb_dt, dim3(1), dim3(256), (const RunGeom*)R->geom, R->rs, (const double*)c->part, R->st,
                  R->st + 1, 1, step_factors(T), 1, nullptr);
    if (rc) return rc;
  }
  if (from <= 1) {
    StageArgs a = stage_args(c);
    for (int q = 0; q < 6; ++q) {
      a.s[q] = c->U[q];
      a.o1[q] = c->F1[q];
      a.o2[q] = c->S[q];
    }
    rc = launch_k(c, 0, k_stage_res<1>, stage_grid(a), dim3(64), a, (const RunState*)R->rs);
    if (rc) return rc;
  }
  if (res && from <= 2) {
    rc = enqueue_cg(c, res_cg(c, L, c->S[4], c->S[5], 0), R->rs, L.spec[0]);
    if (rc) return rc;
  }
  if (res && from <= 3) {
    FinArgs f = fin_args(c, c->S[4], c->S[5], c->S[4], c->S[5], 0.0);
    f.st = R->st;
    rc = launch_k(c, 5, k_finalize1_res, dim3(c->fin_tiles), dim3(kTileX), f, R->rs, c->o.cg_fail_is_error);
    if (rc) return rc;
  }
  if (from <= 4) {
    StageArgs b = stage_args(c);
    for (int q = 0; q < 6; ++q) {
      b.s[q] = c->S[q];
      b.u[q] = c->F1[q];
      b.o1[q] = c->U[q];
    }
    rc = launch_k(c, 1, k_stage_res<2>, stage_grid(b), dim3(64), b, (const RunState*)R->rs);
    if (rc) return rc;
  }
  if (res && from <= 5) {
    rc = enqueue_cg(c, res_cg(c, L, c->U[4], c->U[5], 1), R->rs, L.spec[1]);
    if (rc) return rc;
  }
  if (res && from <= 6) {
    // the final recombination ends the step: step log and, with a CFL step,
    // the CFL partials of the new state (its rho, mx, my, E are final after
    // stage 2; the kernel forms the new B itself)
    FinArgs f = fin_args(c, c->U[4], c->U[5], c->S[4], c->S[5], 0.0);
    f.st = R->st + 1;
    DiagArgs d = diag_args(c, true);
    rc = L.fixed_dt ? launch_k(c, 6, k_finalize2_res<false>, dim3(c->fin_tiles), dim3(kTileX), f, d, R->rs,
                               (const CgState*)R->st, R->log, kLogCap, c->o.cg_fail_is_error)
                    : launch_k(c, 6, k_finalize2_res<true>, dim3(c->fin_tiles), dim3(kTileX), f, d, R->rs,
                               (const CgState*)R->st, R->log, kLogCap, c->o.cg_fail_is_error);
    if (rc) return rc;
    std::swap(c->U[4], c->S[4]);
    std::swap(c->U[5], c->S[5]);
    return MHD_OK;
  }
  DiagArgs d = diag_args(c, true);
  return launch_k(c, 7, k_cfl_res, dim3(L.fixed_dt ? 1 : c->fin_tiles), dim3(kTileX), d, (const RunState*)R->rs,
                  (const CgState*)R->st, R->log, kLogCap, L.fixed_dt ? 0 : 1);
}

// The graph of nsteps (even) steps for the context's current B pointers and
// budgets, captured on first use (least recently used entry replaced).
constexpr int kMaxGraphSteps = 64;   // steps per captured graph (graph = 2)

int step_graph(mhd_ctx* c, Loop& L, int nsteps, StepGraph** out) {
  Resident* R = L.R;
  double* key[4] = {c->U[4], c->U[5], c->S[4], c->S[5]};
  StepGraph* victim = &R->graphs[0];
  for (StepGraph& g : R->graphs) {
    if (g.exec && std::equal(key, key + 4, g.key) && g.spec[0] == L.spec[0] && g.spec[1] == L.spec[1] &&
        g.fixed_dt == L.fixed_dt && g.nsteps == nsteps) {
      g.last_use = ++R->clock;
      *out = &g;
      return MHD_OK;
    }
    if (!g.exec || (victim->exec && g.last_use < victim->last_use)) victim = &g;
  }
  if (victim->exec) (void)hipGraphExecDestroy(victim->exec);
  victim->exec = nullptr;
  HIP_OR_FAIL(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  int rc = MHD_OK;
  for (int k = 0; k < nsteps && !rc; ++k)  // an even number of swaps: the B pointers end where they started
    rc = enqueue_step(c, L, 0);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(c->stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return fail(c, MHD_EHIP, std::string("resident graph capture: ") + hipGetErrorString(e));
  const hipError_t ei = hipGraphInstantiate(&victim->exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ei != hipSuccess) {
    victim->exec = nullptr;
    return fail(c, MHD_EHIP, std::string("resident graph instantiate: ") + hipGetErrorString(ei));
  }
  std::copy(key, key + 4, victim->key);
  victim->spec[0] = L.spec[0];
  victim->spec[1] = L.spec[1];
  victim->nsteps = nsteps;
  victim->fixed_dt = L.fixed_dt;
  victim->last_use = ++R->clock;
  *out = victim;
  return MHD_OK;
}

int read_state(mhd_ctx* c) {
  Resident* R = c->res;
  HIP_OR_FAIL(c, hipMemcpyAsync(R->rs_host, R->rs, sizeof(RunState), hipMemcpyDeviceToHost, c->stream));
  HIP_OR_FAIL(c, hipMemcpyAsync(R->st_host, R->st, 2 * sizeof(CgState), hipMemcpyDeviceToHost, c->stream));
  HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
  return MHD_OK;
}

// After the caller's snapshot callback, which may have used the context
// (mhd_divb reuses the partials buffer; a step swaps the B pointers): take
// the B pointers as they are and recompute the CFL partials of the state.
int rebase(mhd_ctx* c, Loop& L, long long done) {
  for (int k = 0; k < 2; ++k) {
    L.B0[k] = c->U[4 + k];
    L.B1[k] = c->S[4 + k];
  }
  L.base = done;
  // the reference recomputes the primitives of the stored state before the
  // run (integrator_main.cpp:67) and at each snapshot (:107-108)
  if (int rc = reset_raw_density(c)) return rc;
  if (L.fixed_dt) return MHD_OK;
  DiagArgs d = diag_args(c, true);
  return launch_k(c, 7, k_cfl, dim3(c->fin_tiles), dim3(kTileX), d);
}

// A halted step (solve `slot` unfinished after A(spec + 1)): continue the
// solve with host-checked iterations, then enqueue the rest of the step.
int resume_step(mhd_ctx* c, Loop& L, int slot) {
  Resident* R = L.R;
  int rc = launch_k(c, 8, k_clear_halt, dim3(1), dim3(1), R->rs);
  if (rc) return rc;
  const CgArgs a = slot == 0 ? res_cg(c, L, c->S[4], c->S[5], 0) : res_cg(c, L, c->U[4], c->U[5], 1);
  const int nb = c->tiles;
  const bool open = a.open_x | a.open_y;
  auto k_a = open ? k_cg_pass_a_res<true> : k_cg_pass_a_res<false>;
  auto k_b = open ? k_cg_pass_b_res<false, true> : k_cg_pass_b_res<false, false>;
  const long long maxit = R->st_host[slot].max_iters;
  long long it = L.spec[slot] + 1;   // the last A launched
  for (;;) {
    rc = launch_k(c, 4, k_b, dim3(nb), dim3(kTileX), a, (const RunState*)R->rs,
                  (const double*)c->p[it & 1], (int)it);
    if (rc) return rc;
    ++it;
    rc = launch_k(c, 3, k_a, dim3(nb), dim3(kTileX), a, (const RunState*)R->rs,
                  (const double*)c->p[(it - 1) & 1], c->p[it & 1], (int)it);
    if (rc) return rc;
    HIP_OR_FAIL(c, hipMemcpyAsync(R->st_host + slot, R->st + slot, sizeof(CgState), hipMemcpyDeviceToHost, c->stream));
    HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
    if (R->st_host[slot].done) break;
    if (it > maxit + 4) return fail(c, MHD_EHIP, "mhd_run_resident: CG state machine did not terminate");
  }
  if (R->fixed < 0) L.spec[slot] = std::max(L.spec[slot], R->st_host[slot].spmv - 1 + R->margin);
  return enqueue_step(c, L, slot == 0 ? 3 : 6);
}

}  // namespace

extern "C" {

int mhd_resident_tune(mhd_ctx* c, int group_steps, int spec_margin, int fixed_budget, int graph) {
  if (!c) return MHD_EINVAL;
  if (group_steps < 1 || group_steps > kLogCap || spec_margin < 0 || graph < 0 || graph > 2)
    return fail(c, MHD_EINVAL, "mhd_resident_tune: group_steps in [1, 256], spec_margin >= 0, graph 0, 1 or 2");
  int rc = ensure_resident(c);
  if (rc) return rc;
  c->res->group = group_steps;
  c->res->margin = spec_margin;
  c->res->fixed = fixed_budget < 0 ? -1 : fixed_budget;
  c->res->graph = graph;
  return MHD_OK;
}

int mhd_run_resident(mhd_ctx* c, double t_end, double dt_override, int max_steps, int snapshot_every,
                     mhd_snapshot_fn on_snapshot, void* user, int* steps_out, double* t_out, double* step_dt,
                     int* cg_iters) {
  if (!c) return MHD_EINVAL;
  if (!std::isfinite(t_end) || !std::isfinite(dt_override) || dt_override < 0)
    return fail(c, MHD_EINVAL, "mhd_run_resident: bad t_end or dt_override");
  if (snapshot_every < 0) return fail(c, MHD_EINVAL, "mhd_run_resident: snapshot_every must be >= 0");
  int rc = ensure_resident(c);
  if (rc) return rc;
  Resident* R = c->res;
  Loop L{};
  L.R = R;
  for (int k = 0; k < 2; ++k) {
    L.B0[k] = c->U[4 + k];
    L.B1[k] = c->S[4 + k];
  }
  L.fixed_dt = dt_override != 0.0;
  for (int k = 0; k < 2; ++k)
    L.spec[k] = R->fixed >= 0 ? R->fixed : std::max(0, c->last_spmv[k] - 1) + R->margin;

  RunGeom g{};
  g.base = 0;
  g.nx = c->nx;
  g.ny = c->ny;
  g.rows = c->stage_rows;
  g.strips = (c->nx + kStripOut - 1) / kStripOut;
  g.ty = c->tile_rows;
  g.tiles_y = c->tiles_y;
  g.tiles = c->fin_tiles;                    // kb_dt reduces the CFL partials
  g.part_base = 0;
  g.dx = c->dx;
  g.dy = c->dy;
  g.inv_dx = 1.0 / c->dx;
  g.inv_dy = 1.0 / c->dy;
  g.eta = c->eta;
  g.cfl = c->o.cfl;
  g.t_end = t_end;
  g.dt_fixed = dt_override;
  g.max_iters = (c->o.cg_max_iter > 0) ? c->o.cg_max_iter : 2LL * c->n;
  g.max_steps = max_steps < 0 ? -1 : max_steps;
  *R->rs_host = RunState{};
  // the operator of the first dt (integrator_main.cpp:55-60): the CFL step of
  // the state as uploaded, before the primitives are recomputed from the
  // floored density (:67) for the loop's own first CFL step
  double dt_init = dt_override;
  if (dt_override == 0.0 && (rc = mhd_cfl_dt(c, &dt_init))) return rc;
  R->rs_host->prev_dt = dt_init;
  R->rs_host->alpha = (dt_init * c->eta) * c->tab.A_im4;
  HIP_OR_FAIL(c, hipMemcpyAsync(R->geom, &g, sizeof(RunGeom), hipMemcpyHostToDevice, c->stream));
  HIP_OR_FAIL(c, hipMemcpyAsync(R->rs, R->rs_host, sizeof(RunState), hipMemcpyHostToDevice, c->stream));
  HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));   // the host copies above read pageable / pinned host memory

  if (on_snapshot && snapshot_every > 0) {   // initial snapshot (integrator_main.cpp:55-70)
    const double dt0 = dt_init;
    if ((rc = on_snapshot(c, 0, 0.0, dt0, user))) return rc;
  }
  if ((rc = rebase(c, L, 0))) return rc;   // B pointers and the CFL partials of the initial state

  long long done = 0;
  double t = 0.0;
  for (;;) {
    if (max_steps >= 0 && done >= max_steps) break;
    if (!(t < t_end)) break;
    long long G = R->group;
    if (max_steps >= 0) G = std::min<long long>(G, max_steps - done);
    if (snapshot_every > 0) G = std::min<long long>(G, snapshot_every - done % snapshot_every);
    long long k = 0;
    if (R->graph == 2 && c->prof_mask == 0) {   // the group's even part as one graph launch (<= 64 steps each)
      while (k + 2 <= G) {
        const int n = (int)std::min<long long>((G - k) & ~1LL, kMaxGraphSteps);
        StepGraph* sg = nullptr;
        if ((rc = step_graph(c, L, n, &sg))) return rc;
        HIP_OR_FAIL(c, hipGraphLaunch(sg->exec, c->stream));
        k += n;
      }
    } else if (R->graph == 1 && c->prof_mask == 0) {      // pairs of steps (no timing events inside a capture)
      for (; k + 2 <= G; k += 2) {
        StepGraph* sg = nullptr;
        if ((rc = step_graph(c, L, 2, &sg))) return rc;
        HIP_OR_FAIL(c, hipGraphLaunch(sg->exec, c->stream));
      }
    }
    for (; k < G; ++k)
      if ((rc = enqueue_step(c, L, 0))) return rc;
    if ((rc = read_state(c))) return rc;
    // a halted step: finish it on the host, then let the group's remaining steps run again
    while (R->rs_host->halt == kHaltSolve1 || R->rs_host->halt == kHaltSolve2) {
      const int slot = R->rs_host->halt == kHaltSolve1 ? 0 : 1;
      set_parity(c, L, R->rs_host->steps - 1);
      if ((rc = resume_step(c, L, slot))) return rc;
      if ((rc = read_state(c))) return rc;
      const long long in_group = R->rs_host->steps - done;
      if (R->rs_host->halt == 0 && in_group < G && R->rs_host->active && R->rs_host->t < t_end) {
        for (long long k = in_group; k < G; ++k)
          if ((rc = enqueue_step(c, L, 0))) return rc;
        if ((rc = read_state(c))) return rc;
      }
    }
    const RunState& s = *R->rs_host;
    const int halt = s.halt;
    const long long completed = s.steps - ((halt == kHaltCgLimit1 || halt == kHaltCgLimit2) ? 1 : 0);
    set_parity(c, L, completed);
    if (completed > done) {
      HIP_OR_FAIL(c, hipMemcpyAsync(R->log_host, R->log, kLogCap * sizeof(StepLog), hipMemcpyDeviceToHost, c->stream));
      HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
      int most[2] = {0, 0};
      for (long long k = done; k < completed; ++k) {
        const StepLog& e = R->log_host[k % kLogCap];
        if (step_dt) step_dt[k] = e.dt;
        if (cg_iters) {
          cg_iters[2 * k] = e.iters[0];
          cg_iters[2 * k + 1] = e.iters[1];
        }
        for (int q = 0; q < 2; ++q) most[q] = std::max(most[q], e.spmv[q]);
        t = e.t;
      }
      for (int q = 0; q < 2; ++q) {
        if (most[q] > 0) c->last_spmv[q] = most[q];
        if (R->fixed < 0 && most[q] > 0) L.spec[q] = most[q] - 1 + R->margin;
      }
    }
    const bool progressed = completed > done;
    done = completed;
    c->alpha = s.alpha;
    c->have_op = c->have_op || (done > 0);
    if (halt == kHaltNonfinite) return fail(c, MHD_ENONFINITE, "non-finite wave speed (resident loop)");
    if (halt == kHaltCgLimit1 || halt == kHaltCgLimit2) return fail(c, MHD_ECG, "CG reached the iteration limit");
    if (on_snapshot && snapshot_every > 0 && progressed && done % snapshot_every == 0) {
      const double dt_last = R->log_host[(done - 1) % kLogCap].dt;
      if ((rc = on_snapshot(c, (int)done, t, dt_last, user))) return rc;
      if ((rc = rebase(c, L, done))) return rc;
    }
    if (!progressed) break;   // the run had already stopped
  }
  if (steps_out) *steps_out = (int)done;
  if (t_out) *t_out = t;
  return MHD_OK;
}

}  // extern "C"
