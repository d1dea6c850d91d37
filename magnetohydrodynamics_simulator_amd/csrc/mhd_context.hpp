// mhd_context.hpp -- the single-run context (struct mhd_ctx of the C ABI) and
// the host helpers shared by the C-ABI source files (mhd_capi.hip,
// mhd_adjoint.hip).
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <utility>
#include <vector>

#include "../../include/mhd_mi355x.h"
#include "mhd_kernels.hpp"

namespace mhd_host {

constexpr int kNumKernelIds = 9;

struct Tableau {
  // IMEX_Butcher_Tableus, solvers.hpp:39-66
  double gamma = (2.0 - std::sqrt(2.0)) / 2.0;
  double delta = 1.0 - (1.0 / (2.0 * gamma));
  double A_ex3 = gamma, A_ex6 = delta, A_ex7 = 1.0 - delta;
  double A_im4 = gamma, A_im7 = 1.0 - gamma;
  double b_im2 = gamma;
};

struct AdjointStore;   // checkpoints and scratch of the adjoint (mhd_adjoint.hip)
struct Resident;       // driver state and step log of the device-resident loop (mhd_resident.hip)

// Device buffers a step writes besides the state, for the adjoint's recompute:
// the state at the start of the step, the stage state after implicit 1, the
// face B before each implicit solve and the two solutions.
// Any pointer may be null (nothing recorded there).  With use_given set the
// step runs no CG: the two implicit solutions are taken from x1, x2 with the
// final solver states st[0], st[1] (as a previous run of the same step left
// them); otherwise the solutions are copied to x1, x2 and the states to st.
struct StepRecord {
  double* Un[6];
  double* S[6];
  double* SsB[2];
  double* U3sB[2];
  double* x1;
  double* x2;
  int spmv[2];              // products of the two solves (-1: not known on the host)
  int spec[2];              // > 0: run the solve speculatively with this many iterations
  mhd::CgState* st[2];      // device copies of the two solves' final states
  bool use_given;
};

}  // namespace mhd_host

struct mhd_ctx {
  int nx = 0, ny = 0;
  long long n = 0;
  double dx = 0, dy = 0, eta = 0;
  mhd_opts o{};
  mhd_host::Tableau tab;
  hipStream_t stream = nullptr;
  double* U[6] = {};
  double* S[6] = {};
  double* F1[6] = {};
  double* x = nullptr;
  double* r = nullptr;
  // the unfloored density of each cell whose stored density is the floor
  // (the velocities of the reference's last computePrimitives come from it)
  double* rho_raw = nullptr;
  double* p[2] = {};
  double* part_s = nullptr;      // CG start kernel: three partials per tile
  double* part_a = nullptr;      // CG pass A: one partial per tile
  double* part_b = nullptr;      // CG pass B: two partials per tile
  double* part = nullptr;        // diagnostics: two partials per tile
  mhd::CgState* st = nullptr;
  mhd::CgState* st_host = nullptr;    // pinned
  double* part_host = nullptr;   // pinned, two partials per tile
  double* stage = nullptr;       // pinned staging buffer of one field (host transfers)
  double alpha = 0.0;
  bool have_op = false;
  int last_spmv[4] = {1, 1, 1, 1};   // products of the previous solve per slot (0, 1: the step's
                                      // solves; 2, 3: the adjoint's)
  int tiles = 1, tiles_y = 1, tile_rows = 4;   // 2D sweeps: kTileX columns x tile_rows rows per tile
  // the recombinations and the CFL partials: about kFinTargetTiles tiles (their
  // partials are a max and a count, exact in any grouping)
  int fin_tiles = 1, fin_tiles_y = 1, fin_rows = 4;
  int stage_rows = 16;
  std::string err;
  unsigned prof_mask = 0;        // kernel ids whose launches are bracketed by events
  long long prof_n[mhd_host::kNumKernelIds] = {};
  double prof_ms[mhd_host::kNumKernelIds] = {};
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pending[mhd_host::kNumKernelIds];
  mhd_host::AdjointStore* adj = nullptr;
  mhd_host::Resident* res = nullptr;
};

namespace mhd_host {

int fail(mhd_ctx* c, int code, const std::string& msg);

#define HIP_OR_FAIL(ctx, call)                                                                  \
  do {                                                                                          \
    hipError_t e_ = (call);                                                                     \
    if (e_ != hipSuccess)                                                                       \
      return ::mhd_host::fail((ctx), MHD_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Launch helper: optional event bracketing for per-kernel timing.
template <typename F>
int launch(mhd_ctx* c, int id, F&& body) {
  hipEvent_t a = nullptr, b = nullptr;
  const bool prof = (c->prof_mask >> id) & 1u;
  if (prof) {
    HIP_OR_FAIL(c, hipEventCreate(&a));
    HIP_OR_FAIL(c, hipEventCreate(&b));
    HIP_OR_FAIL(c, hipEventRecord(a, c->stream));
  }
  body();
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(c, MHD_EHIP, std::string("kernel launch: ") + hipGetErrorString(e));
  if (prof) {
    HIP_OR_FAIL(c, hipEventRecord(b, c->stream));
    c->pending[id].emplace_back(a, b);
  }
  return MHD_OK;
}

// Launch of one kernel of kernel id `id` on the context's stream.  With
// timing on for that id, the kernel's own dispatch records the start and end
// timestamps (hipExtLaunchKernelGGL's events): no separate event markers go
// into the stream, so timing the hot kernels inside a timed region leaves the
// step's launch sequence as it is.
template <typename K, typename... Args>
int launch_k(mhd_ctx* c, int id, K kernel, dim3 grid, dim3 block, Args... args) {
  const bool prof = (c->prof_mask >> id) & 1u;
  if (prof) {
    hipEvent_t a = nullptr, b = nullptr;
    HIP_OR_FAIL(c, hipEventCreate(&a));
    HIP_OR_FAIL(c, hipEventCreate(&b));
    hipExtLaunchKernelGGL(kernel, grid, block, 0, c->stream, a, b, 0, args...);
    c->pending[id].emplace_back(a, b);
  } else {
    hipLaunchKernelGGL(kernel, grid, block, 0, c->stream, args...);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(c, MHD_EHIP, std::string("kernel launch: ") + hipGetErrorString(e));
  return MHD_OK;
}

int grid_1d(long long n);
// Kernel argument blocks of the context's current buffers (coefficients from
// the host-side dt / operator; the resident loop's kernels replace them with
// the device driver state's).
mhd::StageArgs stage_args(mhd_ctx* c);
// the launch grid of a stage kernel with these arguments (StageArgs.band)
dim3 stage_grid(const mhd::StageArgs& a);
mhd::CgArgs cg_args(mhd_ctx* c, const double* bx, const double* by);
mhd::FinArgs fin_args(mhd_ctx* c, const double* bx, const double* by, double* bxo, double* byo, double coef);
mhd::DiagArgs diag_args(mhd_ctx* c, bool fin);
// One CG solve (Eigen's algorithm) of A x = eta J(bx, by) or A x = bin into
// c->x.  spec == 0: chunks predicted from the slot's previous solve, the host
// reads the state after each chunk until the solve is done, *out receives it.
// spec > 0: spec iterations are launched without any host round trip (the
// passes of a finished solve return at once), the final device state is
// copied to spec_out, and out->done is set to -1 (not known).
int cg_run(mhd_ctx* c, const double* bx, const double* by, const double* bin, mhd::CgState* out, int slot,
           int spec = 0, mhd::CgState* spec_out = nullptr);
int to_host_padded(mhd_ctx* c, const double* dev, double* host, bool ghosts);
// Ghost layers of one padded (ny+4) x (nx+4) host array: periodic
// (Grid::fillOneField) or, on an outflow axis, the nearest inside value.
void fill_ghosts(double* f, int nx, int ny, int open_x, int open_y);
// A computePrimitives pass over the stored state (Grid.hpp:111-115) as far as
// the next CFL step sees it: the velocities of floored cells come from the
// floored density from now on (k_raw_reset).
int reset_raw_density(mhd_ctx* c);
// dst[q] = src[q] for q < n (n <= 6 device fields of the context's size), one launch.
int copy_fields(mhd_ctx* c, double* const* dst, double* const* src, int n);
int from_host_padded(mhd_ctx* c, const double* host, double* dev);
// One IMEX step (mhd_imex_step); rec, if not null, receives the step's data.
int imex_step(mhd_ctx* c, double dt, mhd_step_stats* stats, StepRecord* rec);
void destroy_adjoint(mhd_ctx* c);
void destroy_resident(mhd_ctx* c);

}  // namespace mhd_host
