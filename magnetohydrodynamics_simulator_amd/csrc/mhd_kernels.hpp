// mhd_kernels.hpp -- kernel argument blocks and launch constants shared by the
// kernels (mhd_kernels.hip) and the C-ABI host code (mhd_capi.hip).
#pragma once

#include <hip/hip_runtime.h>

namespace mhd {

constexpr int kStripOut = 60;     // output columns per 64-lane strip of the stage kernel
constexpr int kTileX = 256;       // columns (= threads) per tile of the 2D sweeps
#ifndef MHD_TARGET_TILES
#define MHD_TARGET_TILES 2048
#endif
constexpr int kTargetTiles = MHD_TARGET_TILES;   // tiles per 2D sweep (compile-time tuning parameter)

// Explicit-stage kernel arguments.  MODE 0: o1 = L(s).
// MODE 1 (f1 = L(s)): o1 = W = s + c_b f1, o2 = U2* = s + c_a f1.
// MODE 2 (f2 = L(s)): o1 = U3* = u + c_b f2 [+ c_g g2(s) on B], u = W.
struct StageArgs {
  const double* s[6];   // state whose fluxes are computed (halo reads)
  const double* u[6];   // MODE 2: W
  double* o1[6];
  double* o2[6];
  int nx, ny, rows;
  double dx, dy, inv_dx, inv_dy, eta;
  double c_a, c_b, c_g;
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
  double dx, dy, eta;
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
  double dx, dy, eta, c;
  const double* x;
  const double* bx;
  const double* by;
  double* bx_out;
  double* by_out;
  const CgState* st;
};

struct DiagArgs {
  int nx, ny, ty, tiles_y;
  double dx, dy;
  const double* s[6];
  double* part;    // two partials per tile
};

template <int MODE> __global__ void k_stage(StageArgs a);
template <bool FROM_B> __global__ void k_cg_start(CgArgs a, const double* bin);
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
__global__ void k_copy(const double* src, double* dst, long long n);

}  // namespace mhd
