// mhd_kernels.hpp -- kernel argument blocks and launch constants shared by the
// kernels (mhd_kernels.hip) and the C-ABI host code (mhd_capi.hip).
#pragma once

#include <hip/hip_runtime.h>

namespace mhd {

constexpr int kStripOut = 60;     // output columns per 64-lane strip of the stage kernel
#ifndef MHD_STAGE_BANDS
#define MHD_STAGE_BANDS 1          // column bands of the single-run stage launch (0: a 2D grid; tuning parameter)
#endif
constexpr int kStageBands = MHD_STAGE_BANDS;
constexpr int kTileX = 256;       // columns (= threads) per tile of the 2D sweeps
#ifndef MHD_TARGET_TILES
#define MHD_TARGET_TILES 2048
#endif
constexpr int kTargetTiles = MHD_TARGET_TILES;   // tiles per 2D sweep (compile-time tuning parameter)
#ifndef MHD_FIN_TARGET_TILES
#define MHD_FIN_TARGET_TILES 1024
#endif
constexpr int kFinTargetTiles = MHD_FIN_TARGET_TILES;   // tiles of the recombinations and CFL partials

// Explicit-stage kernel arguments.  MODE 0: o1 = L(s).
// MODE 1 (f1 = L(s)): o1 = W = s + c_b f1, o2 = U2* = s + c_a f1.
// MODE 2 (f2 = L(s)): o1 = U3* = u + c_b f2 [+ c_g g2(s) on B], u = W.
struct StageArgs {
  const double* s[6];   // state whose fluxes are computed (halo reads)
  const double* u[6];   // MODE 2: W
  double* o1[6];
  double* o2[6];
  int nx, ny, rows;
  int open_x, open_y;   // outflow (zero-gradient) boundaries instead of periodic ones
  double dx, dy, inv_dx, inv_dy, eta;
  double c_a, c_b, c_g;
  // MODE 2: the new density before its floor, where it is <= 1e-8 (null: not kept)
  double* rho_raw;
  // launch shape: band == 0: a (strips, segs) grid; band > 0: a 1D grid of
  // strips * segs workgroups, each XCD taking a contiguous run of the
  // column bands of `band` strips (rows of strips bottom to top in a band)
  int strips, segs, band;
};

// Conjugate-gradient state, device resident.  Written by block 0 of a pass
// kernel, read by later kernels; abs_new is double-buffered by iteration parity.
struct CgState {
  double rhs_norm2, threshold, abs_new[2], alpha, beta, rn2, pap;
  long long max_iters;
  int iters, spmv, done;   // done: 0 running, 1 converged, 2 max iterations, 3 zero rhs
};

struct CgArgs {
  int nx, ny, ty, tiles_y;
  int open_x, open_y;
  double dx, dy, eta;
  double inv_dx, inv_dy;   // 1 / dx, 1 / dy (the right-hand side eta J(B))
  double invd, diag, cx, cy, tol;
  const double* bx;  // face B for the right-hand side
  const double* by;
  double* r;
  double* p0;
  double* x;
  double* part_s;  // start kernel: three partials per tile (|b|^2, b . p, p . Ap)
  double* part_a;  // pass A: one partial per tile (p . Ap)
  double* part_b;  // pass B: two partials per tile (|r|^2, r . z)
  CgState* st;
};

struct FinArgs {
  int nx, ny, ty, tiles_y;
  int open_x, open_y;
  double dx, dy, eta, c;
  double inv_dx, inv_dy;   // 1 / dx, 1 / dy: the differences of the curl and J multiply by them
  const double* x;
  const double* bx;
  const double* by;
  double* bx_out;
  double* by_out;
  const CgState* st;
};

struct DiagArgs {
  int nx, ny, ty, tiles_y;
  int open_x, open_y;
  double dx, dy;
  const double* s[6];
  double* part;    // two partials per tile
  // the density a floored cell had before its floor (null: the stored one);
  // read only where the stored density is the floor itself
  const double* rho_raw;
};

// ---------------------------------------------------------------------------
// Batched ensembles: many independent periodic runs of different sizes share
// pooled field arrays (run r owns cells [base_r, base_r + nx_r ny_r) of every
// field) and advance together, one launch per kernel for all runs.  A block
// finds its run and its block index within the run in a job table.
// ---------------------------------------------------------------------------

// Geometry, layout and schedule of one run (host-written at creation).
struct RunGeom {
  long long base;          // first cell of the run in the pooled arrays
  int nx, ny;
  int rows, strips;        // stage kernel: rows per segment, 60-column strips
  int ty, tiles_y, tiles;  // 2D sweeps: tile rows, tiles per column of tiles, tiles
  int part_base;           // first tile of the run in the pooled partial arrays
  double dx, dy, inv_dx, inv_dy, eta;
  double cfl;              // effective Courant number
  double t_end;            // the run stops once t >= t_end
  double dt_fixed;         // > 0: constant time step (the driver's override)
  long long max_iters;     // CG iteration limit
  int max_steps;           // < 0: no step limit
  int open;                // outflow axes: bit 0 x, bit 1 y (the batch's mhd_opts.bc_x / bc_y)
};

// Driver state of one run, device resident; kb_dt writes it once per step.
// halt (single-run resident loop only): a kHalt* code once the loop has
// stopped inside a step; every kernel of the loop then returns at once.
// coef: the step's stage coefficients dt * (A_ex[3], A_ex[6], A_ex[7],
// A_im[7], A_im[4]), read by the stage kernels and finalisers (loaded from
// memory they live in scalar registers; formed in a stage kernel they would
// cost it vector registers it does not have).
struct RunState {
  double t, dt, prev_dt, alpha, cx, cy, diag, invd;
  double coef[5];
  int steps, active, nonfinite, halt;
};

// The tableau weights kb_dt multiplies by dt (IMEX_Butcher_Tableus, solvers.hpp:39-66).
struct StepFactors {
  double ex3, ex6, ex7, im7, im4;
};
template <typename Tab>
inline StepFactors step_factors(const Tab& T) {
  return StepFactors{T.A_ex3, T.A_ex6, T.A_ex7, T.A_im7, T.A_im4};
}

constexpr int kHaltSolve1 = 1;      // implicit solve 1 not finished within the launched iterations
constexpr int kHaltSolve2 = 2;      // implicit solve 2 likewise
constexpr int kHaltNonfinite = 3;   // non-finite wave speed: the step was not taken
constexpr int kHaltCgLimit1 = 4;    // solve 1 reached the iteration limit (opts.cg_fail_is_error)
constexpr int kHaltCgLimit2 = 5;    // solve 2 likewise

// One entry of the resident loop's step log (ring buffer on the device).
struct StepLog {
  double dt, t, err[2];
  int iters[2], spmv[2];
};

struct BatchStageArgs {
  const RunGeom* geom;
  const RunState* rs;
  const int2* jobs;        // (run, wave index within the run)
  const double* s[6];
  const double* u[6];
  double* o1[6];
  double* o2[6];
  const int* ctl;          // batch control word (kCtlHalt, kCtlSteps); null: none
  double* rho_raw;         // pooled: MODE 2 keeps the unfloored density of floored cells (null: none)
};

struct BatchCgArgs {
  const RunGeom* geom;
  const RunState* rs;
  const int2* jobs;        // (run, tile index within the run)
  const double* bx;
  const double* by;
  double* r;
  double* p0;
  double* x;
  double* part_s;
  double* part_a;
  double* part_b;
  CgState* st;             // one state per run for this solve slot
  double tol;
  const int* ctl;
};

struct BatchFinArgs {
  const RunGeom* geom;
  const RunState* rs;
  const int2* jobs;
  const double* x;
  const double* bx;
  const double* by;
  double* bx_out;
  double* by_out;
  const CgState* st;
  const int* ctl;
};

struct BatchDiagArgs {
  const RunGeom* geom;
  const RunState* rs;
  const int2* jobs;
  const double* s[6];
  double* part;
  const int* ctl;
  const double* rho_raw;   // pooled (null: none)
};

// Control word of a batch's device-resident group of steps (an int[4] on the
// device): [0] halt code (0 running; kHaltSolve1 / kHaltSolve2: a solve left
// unfinished after its launched iterations; kHaltNonfinite; kHaltCgLimit1 /
// 2: a solve at its iteration limit with opts.cg_fail_is_error), [1] batch
// steps begun since the group started (kb_dt), [2], [3] unused.  Once [0] is
// set every batched kernel returns at once until the host clears it.
constexpr int kCtlHalt = 0;
constexpr int kCtlSteps = 1;

template <int MODE> __global__ void kb_stage(BatchStageArgs b);
// OPEN: the batch has an outflow axis (RunGeom.open); the kernels are built
// per boundary mode like the single-run ones.
template <bool OPEN> __global__ void kb_cg_start(BatchCgArgs b);
template <bool OPEN> __global__ void kb_cg_start_rhs(BatchCgArgs b, const double* bin);
template <bool OPEN> __global__ void kb_cg_pass_a(BatchCgArgs b, const double* pold, double* pnew, int it);
template <bool FIRST, bool OPEN> __global__ void kb_cg_pass_b(BatchCgArgs b, const double* pv, int it);
__global__ void kb_finalize1(BatchFinArgs b);
template <bool OPEN> __global__ void kb_finalize2(BatchFinArgs b);
__global__ void kb_cfl(BatchDiagArgs b);
template <bool OPEN> __global__ void kb_finalize2_cfl(BatchFinArgs b, BatchDiagArgs d);
__global__ void kb_dt(const RunGeom* geom, RunState* rs, const double* part, CgState* st0, CgState* st1,
                      int nruns, StepFactors f, int halt_on_nonfinite, int* ctl);
// After a batch's solve `slot` (0, 1): halts the group (ctl[kCtlHalt]) if a
// stepping run's solve is unfinished, or at its limit with fail_is_error;
// records each stepping run's iteration count in iters[2 run + slot].
__global__ void kb_check_solve(const RunState* rs, const CgState* st, int nruns, int slot, int fail_is_error,
                               int* ctl, int* iters);
__global__ void kb_ctl_reset(int* ctl);

// Single-run resident loop: the plain kernels with their step coefficients
// read from the device driver state (c = dt * f), returning at once once the
// run has stopped or the loop has halted.
template <int MODE> __global__ void k_stage_res(StageArgs a, const RunState* rs);
template <bool OPEN> __global__ void k_cg_start_res(CgArgs a, const RunState* rs);
template <bool OPEN>
__global__ void k_cg_pass_a_res(CgArgs a, const RunState* rs, const double* pold, double* pnew, int it);
template <bool FIRST, bool OPEN>
__global__ void k_cg_pass_b_res(CgArgs a, const RunState* rs, const double* pv, int it);
__global__ void k_finalize1_res(FinArgs a, RunState* rs, int fail_is_error);
template <bool CFL>
__global__ void k_finalize2_res(FinArgs a, DiagArgs d, RunState* rs, const CgState* st, StepLog* log, int log_cap,
                                int fail_is_error);
__global__ void k_cfl_res(DiagArgs a, const RunState* rs, const CgState* st, StepLog* log, int log_cap,
                          int partials);
__global__ void k_clear_halt(RunState* rs);

template <int MODE> __global__ void k_stage(StageArgs a);
template <bool FROM_B> __global__ void k_cg_start(CgArgs a, const double* bin);
template <bool FROM_B> __global__ void k_cg_start_open(CgArgs a, const double* bin);
__global__ void k_cg_pass_a_open(CgArgs a, const double* pold, double* pnew, int it);
template <bool FIRST> __global__ void k_cg_pass_b_open(CgArgs a, const double* pv, int it);
__global__ void k_cg_rhs(CgArgs a);
__global__ void k_cg_pass_a(CgArgs a, const double* pold, double* pnew, int it);
template <bool FIRST> __global__ void k_cg_pass_b(CgArgs a, const double* pv, int it);
__global__ void k_finalize1(FinArgs a);
__global__ void k_finalize2(FinArgs a);
__global__ void k_rate(FinArgs a, int zero);
__global__ void k_cfl(DiagArgs a);
__global__ void k_divb(DiagArgs a);
__global__ void k_prims(DiagArgs a, double* vx, double* vy, double* p, double* bxc, double* byc);
__global__ void k_floor_rho(double* rho, int n);
// The density floor of an uploaded state, keeping the unfloored value of each
// floored cell in raw (Grid::computePrimitives forms the velocities from it).
__global__ void k_floor_rho_raw(double* rho, double* raw, int n);
// raw = rho where rho is at the floor: the velocities of the next CFL step come
// from the floored density (a computePrimitives pass over the stored state).
__global__ void k_raw_reset(const double* rho, double* raw, int n);
__global__ void k_cg_reset(CgState* st, long long max_iters);
struct CopyArgs {
  const double* src[6];
  double* dst[6];
  int nf;
};
__global__ void k_copy_fields(CopyArgs a, long long n);
__global__ void k_copy(const double* src, double* dst, long long n);
struct MixArgs {
  const double* in[12];
  double* out[12];
};
template <int NI, int NO> __global__ void k_stream_mix(MixArgs a, long long n);

}  // namespace mhd
