// mhd_batch.hpp -- the batch object behind the mhd_batch_* C ABI and the host
// helpers that the batched adjoint (mhd_batch_adjoint.hip) shares with the
// batched step (mhd_batch.hip).  Internal to the library.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

#include "../../include/mhd_mi355x.h"
#include "mhd_kernels.hpp"

namespace mhd_batch_host {

struct BatchTableau {
  double gamma = (2.0 - std::sqrt(2.0)) / 2.0;
  double delta = 1.0 - (1.0 / (2.0 * gamma));
  double A_ex3 = gamma, A_ex6 = delta, A_ex7 = 1.0 - delta;
  double A_im4 = gamma, A_im7 = 1.0 - gamma;
};

struct BatchAdjoint;   // mhd_batch_adjoint.hip

// An even number of batch steps (a pair, or a whole group) captured as a
// graph, valid for the step count, B pointers, CG budgets and CFL fusion it
// was captured with.
struct BatchGraph {
  hipGraphExec_t exec = nullptr;
  double* key[4] = {};
  int spec[2] = {0, 0};
  int nsteps = 0;
  bool fused = true;
  unsigned long long last_use = 0;
};
constexpr int kBatchGraphCache = 4;

}  // namespace mhd_batch_host

struct mhd_batch {
  int nruns = 0;
  long long total = 0;              // cells of all runs
  int total_tiles = 0;
  mhd_opts o{};
  mhd_batch_host::BatchTableau tab;
  hipStream_t stream = nullptr;
  std::vector<mhd::RunGeom> geom_h;
  mhd::RunGeom* geom = nullptr;
  mhd::RunState* rs = nullptr;
  mhd::RunState* rs_host = nullptr; // pinned
  double* U[6] = {};
  double* S[6] = {};
  double* W[6] = {};
  double* x = nullptr;
  double* r = nullptr;
  double* rho_raw = nullptr;        // pooled: the unfloored density of floored cells (DiagArgs.rho_raw)
  double* p[2] = {};
  double* part = nullptr;           // CFL: two partials per tile
  double* part_s = nullptr;
  double* part_a = nullptr;
  double* part_b = nullptr;
  mhd::CgState* st[2] = {};
  mhd::CgState* st_host = nullptr;  // pinned, one solve slot (nruns states)
  int2* stage_jobs = nullptr;
  int n_stage_jobs = 0;
  int2* tile_jobs = nullptr;
  int n_tile_jobs = 0;
  size_t cap_stage_jobs = 0, cap_tiles = 0;   // capacity of the job tables and partial arrays
  int stage_rows = 32, tile_rows = 32;
  int last_spmv[3] = {1, 1, 1};     // longest solve of the last call per slot (0, 1: step, 2: adjoint)
  bool any_eta = false;
  // part holds the CFL partials of the current state (formed by the last
  // step's final recombination); cleared by anything that changes the state,
  // the schedule or the tile plan
  bool cfl_fresh = false;
  long long steps_taken = 0;
  std::vector<int> cg_iters_last;   // 2 per run: Eigen iteration counts of the last step
  double* stage = nullptr;          // pinned staging buffer of the largest run (host transfers)
  // device-resident groups of steps (mhd_batch_run; mhd_batch_tune)
  int group = 16;                   // steps enqueued between two read-backs (0: one host-checked step at a time)
  int margin = 0;                   // extra launched CG iterations per solve
  int graph = 0;                    // 1: replay pairs of steps as captured graphs; 2: each group as one graph
  int fused_cfl = 1;                // 1: the final recombination forms the next CFL partials; 0: a separate pass
  int* ctl = nullptr;               // device control word (kCtlHalt, kCtlSteps)
  int* ctl_host = nullptr;          // pinned
  int* iters = nullptr;             // device: 2 per run, the last step's Eigen iteration counts
  int* iters_host = nullptr;        // pinned
  mhd::CgState* st_host2 = nullptr; // pinned, the other solve slot
  mhd_batch_host::BatchGraph graphs[mhd_batch_host::kBatchGraphCache];
  unsigned long long gclock = 0;
  mhd_batch_host::BatchAdjoint* adj = nullptr;
  std::string err;
};

namespace mhd_batch_host {

int bfail(mhd_batch* b, int code, const std::string& msg);

#define BHIP(b, call)                                                                           \
  do {                                                                                          \
    hipError_t e_ = (call);                                                                     \
    if (e_ != hipSuccess)                                                                       \
      return mhd_batch_host::bfail((b), MHD_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define BLAUNCH(b, ...)                                                                         \
  do {                                                                                          \
    hipLaunchKernelGGL(__VA_ARGS__);                                                            \
    hipError_t e_ = hipGetLastError();                                                          \
    if (e_ != hipSuccess)                                                                       \
      return mhd_batch_host::bfail((b), MHD_EHIP, std::string("kernel launch: ") + hipGetErrorString(e_)); \
  } while (0)

// One CG solve of every run whose state in `st` is not done (Eigen's
// algorithm per run).  The right-hand side is eta J(bx, by) or, when bin is
// set, bin; the solution goes to x, the final states to st (nruns entries),
// the operator comes from rs.  spec > 0: spec iterations with no host check
// (a run stops at its own convergence; the caller verifies the states);
// spec == 0: the host checks all runs after each chunk, the first chunk as
// long as the longest solve of `slot` in the last call.
struct CgCall {
  int slot = 0;
  const double* bx = nullptr;
  const double* by = nullptr;
  const double* bin = nullptr;
  double* x = nullptr;
  mhd::CgState* st = nullptr;
  const mhd::RunState* rs = nullptr;
  int spec = 0;
};
int cg_batch(mhd_batch* b, const CgCall& call, int* most_spmv = nullptr);

mhd::BatchFinArgs fin_args(mhd_batch* b, const mhd::RunState* rs, const double* x, const mhd::CgState* st,
                           const double* bx, const double* by, double* bxo, double* byo);

void fill_ghosts_padded(double* f, int nx, int ny, int open_x, int open_y);

// The batch's boundary mode (all runs share it) selects the kernels built for
// it: periodic, or with an outflow axis (the `OPEN` variants).
inline bool batch_open(const mhd_batch* b) { return b->o.bc_x == MHD_BC_OUTFLOW || b->o.bc_y == MHD_BC_OUTFLOW; }
inline auto kb_start(const mhd_batch* b) { return batch_open(b) ? mhd::kb_cg_start<true> : mhd::kb_cg_start<false>; }
inline auto kb_start_rhs(const mhd_batch* b) { return batch_open(b) ? mhd::kb_cg_start_rhs<true> : mhd::kb_cg_start_rhs<false>; }
inline auto kb_pass_a(const mhd_batch* b) { return batch_open(b) ? mhd::kb_cg_pass_a<true> : mhd::kb_cg_pass_a<false>; }
template <bool FIRST>
inline auto kb_pass_b(const mhd_batch* b) { return batch_open(b) ? mhd::kb_cg_pass_b<FIRST, true> : mhd::kb_cg_pass_b<FIRST, false>; }
inline auto kb_fin2(const mhd_batch* b) { return batch_open(b) ? mhd::kb_finalize2<true> : mhd::kb_finalize2<false>; }
inline auto kb_fin2_cfl(const mhd_batch* b) { return batch_open(b) ? mhd::kb_finalize2_cfl<true> : mhd::kb_finalize2_cfl<false>; }
int stage_in(mhd_batch* b, int run, const double* host, double* pooled);
int stage_out(mhd_batch* b, int run, const double* pooled, double* host);

void destroy_batch_adjoint(mhd_batch* b);

}  // namespace mhd_batch_host
