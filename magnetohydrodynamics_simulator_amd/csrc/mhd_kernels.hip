// mhd_kernels.hip -- gfx950 kernels of the 2D resistive-MHD ARS(2,2,2) step.
//
// Device layout: every field is a compact, periodic nx * ny array of doubles,
// row-major (index j * nx + i), structure of arrays.  bx[j*nx+i] is B_x on the
// left face x_{i-1/2} of cell (i, j), by[j*nx+i] is B_y on the bottom face
// y_{j-1/2}, corner quantities (E_z, the CG unknowns) sit at the bottom-left
// corner -- the staggering of the reference's Grid (Grid.hpp:19-26,
// solvers.hpp:312), without its ghost layers: periodicity is applied in the
// index arithmetic instead of by ghost-fill passes.
#include <hip/hip_runtime.h>

#include "mhd_device.hpp"
#include "mhd_kernels.hpp"

namespace mhd {

// ---------------------------------------------------------------------------
// Fused explicit stage: primitives -> MUSCL -> HLL x/y -> upwind CT E_z ->
// flux divergence and CT face update -> stage combination, in one pass.
//
// One wavefront owns a strip of 64 columns, lane l <-> column x0 - 2 + l, and
// marches up a segment of rows.  Lanes 2..61 produce output (60 columns per
// strip); lanes 0,1,62,63 are the halo the stencil needs.  Row k of the march
// computes, with everything older kept in registers:
//   P(k)       primitives of row k                     (loads of row k)
//   YR(k-1)    y-reconstruction of row k-1             (P(k-2), P(k-1), P(k))
//   G(k-2)     y-face flux on top of row k-2           (YR(k-2).R, YR(k-1).L)
//   XR(k-1)    x-reconstruction of row k-1             (cross-lane neighbours)
//   F(k-1)     x-face flux right of row k-1 cells      (XR of this and next lane)
//   Ez(k-1)    corner E_z, bottom-left of row k-1      (F(k-1), F(k-2), G(k-2))
//   dU(k-2)    flux divergence + CT of row k-2         (F, G, Ez)
// so output row k-2 is final once row k has been loaded; a segment of H rows
// takes H + 4 row steps.  Neighbouring columns move between lanes with
// cross-lane shuffles; no barriers are used (a workgroup is one wavefront).
// Stage 1 keeps each row's point values for its output two rows later in a
// per-wavefront LDS ring instead of re-reading them from memory.
// ---------------------------------------------------------------------------

struct Carry {
  // primitives of rows k-2 (a) and k-1 (b): rho (floored), vx, vy, p, bxc, byc
  double a_rho, a_vx, a_vy, a_p, a_bxc;
  double b_rho, b_vx, b_vy, b_p, b_bxc, b_byc;
  double yr_rho, yr_vx, yr_vy, yr_p, yr_bx;   // YR(k-2).R
  double f_rho, f_mx, f_my, f_E;              // F(c, k-2)
  double fby_left;                            // F.by(c-1, k-2)
  double g_rho, g_mx, g_my, g_E;              // G(c, k-3)
  double ez;                                  // Ez(c, k-2)
  double vx_left_a, vy_left_a;                // vx, vy of (c-1, k-2)
  double bxr_b;                               // bxf(c+1, k-1)
  double byf_b;                               // byf(c, k-1)
  double byf_c;                               // byf(c, k)
  // resistive explicit term of stage 2: J at (c, k-2) and face B history
  double j_a;
  double bxl_b, bxl_a;                        // bxf(c, k-1), bxf(c, k-2)
};

// One-lane shifts across the whole 64-lane wavefront as DPP moves (wave_shr:1,
// wave_shl:1) on the two halves of the double: VALU moves instead of LDS
// permutes.  The lane without a source (lane 0 up, lane 63 down) receives 0
// (bound_ctrl); those lanes are stencil halo, and what they compute from it
// reaches only halo lanes (lanes 0-1 and 62-63 of a strip store nothing).
// A plain DPP move has no previous destination value to preserve, so it needs
// no copy of the source into the destination first (32 fewer VALU
// instructions per row of the march; the stored results are bitwise the same).
constexpr int kDppWaveShl1 = 0x130, kDppWaveShr1 = 0x138;
template <int CTRL>
__device__ __forceinline__ double wave_shift(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xf, 0xf, true);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double sh_up(double v) { return wave_shift<kDppWaveShr1>(v); }   // value of lane l-1
__device__ __forceinline__ double sh_dn(double v) { return wave_shift<kDppWaveShl1>(v); }   // value of lane l+1

// One wavefront's work: strip `strip` (columns strip * 60 ...), segment `seg`
// (rows seg * a.rows ...).
template <int MODE>
__device__ __forceinline__ void stage_body(const StageArgs& a, int strip, int seg) {
  const int lane = threadIdx.x;
  const int nx = a.nx, ny = a.ny;
  const int x0 = strip * kStripOut;
  const int y0 = seg * a.rows;
  const int y1 = min(y0 + a.rows, ny);
  const int col = x0 - 2 + lane;
  const int ox = a.open_x, oy = a.open_y;
  const int c0 = bc_mod(col, nx, ox);
  const int c1 = bc_mod(col + 1, nx, ox);
  const bool out_lane = (lane >= 2) && (lane < 2 + kStripOut) && (col < nx);
  const double idx_ = a.inv_dx, idy_ = a.inv_dy;

  Carry s = {};
  s.byf_c = a.s[5][bc_wrap(y0 - 2, ny, oy) * nx + c0];
  // stage 1: the point values of row k (rho, mx, my, E, bx and by faces) are
  // kept for two rows in a per-wave LDS ring instead of being re-read from
  // memory when row k is output
  __shared__ double ring[MODE == 1 ? 3 : 1][6][64];
  int slot = 0;

  s.bxl_b = 0.0;
  // row loads are issued one row ahead of their use so that their latency
  // overlaps the flux arithmetic of the current row
  double n_rho, n_mx, n_my, n_E, n_bxl, n_bxr, n_byn;
  {
    const int b0 = bc_wrap(y0 - 2, ny, oy) * nx;
    const int b1 = bc_wrap(y0 - 1, ny, oy) * nx;
    n_rho = a.s[0][b0 + c0]; n_mx = a.s[1][b0 + c0]; n_my = a.s[2][b0 + c0]; n_E = a.s[3][b0 + c0];
    n_bxl = a.s[4][b0 + c0]; n_bxr = a.s[4][b0 + c1]; n_byn = a.s[5][b1 + c0];
  }

  for (int k = y0 - 2; k <= y1 + 1; ++k) {
    // ---- loads of row k (issued last iteration) and of row k + 1 ----
    const double rho = n_rho, mx = n_mx, my = n_my, E = n_E, bxl = n_bxl, bxr = n_bxr, byn = n_byn;
    {
      // unconditional (rows past the segment are valid periodic rows; the last
      // iteration's values are unused) so the wait counters stay static
      const int bn = bc_wrap(k + 1, ny, oy) * nx;
      const int bnn = bc_wrap(k + 2, ny, oy) * nx;
      n_rho = a.s[0][bn + c0]; n_mx = a.s[1][bn + c0]; n_my = a.s[2][bn + c0]; n_E = a.s[3][bn + c0];
      n_bxl = a.s[4][bn + c0]; n_bxr = a.s[4][bn + c1]; n_byn = a.s[5][bnn + c0];
    }
    if (MODE == 1) {
      ring[slot][0][lane] = rho; ring[slot][1][lane] = mx; ring[slot][2][lane] = my;
      ring[slot][3][lane] = E; ring[slot][4][lane] = bxl; ring[slot][5][lane] = s.byf_c;
    }
    // ---- P(k) ----
    const double bxc = (bxl + bxr) / 2.0;

    const double byc = (s.byf_c + byn) / 2.0;
    double c_rho, c_vx, c_vy, c_p;
    primitives(rho, mx, my, E, bxc, byc, c_rho, c_vx, c_vy, c_p);

    // ---- YR(k-1) and G(k-2) ----
    double yl_rho, yr_rho, yl_vx, yr_vx, yl_vy, yr_vy, yl_p, yr_p, yl_bx, yr_bx;
    muscl(s.a_rho, s.b_rho, c_rho, yl_rho, yr_rho);
    muscl(s.a_vx, s.b_vx, c_vx, yl_vx, yr_vx);
    muscl(s.a_vy, s.b_vy, c_vy, yl_vy, yr_vy);
    muscl(s.a_p, s.b_p, c_p, yl_p, yr_p);
    muscl(s.a_bxc, s.b_bxc, bxc, yl_bx, yr_bx);
    Flux G;
    {
      Prim D{s.yr_rho, s.yr_vx, s.yr_vy, s.yr_p, s.yr_bx, s.byf_b};
      Prim U{yl_rho, yl_vx, yl_vy, yl_p, yl_bx, s.byf_b};
      G = hll_y(D, U);
    }

    // ---- XR(k-1) and F(k-1) ----
    const double l_rho = sh_up(s.b_rho), r_rho = sh_dn(s.b_rho);
    const double l_vx = sh_up(s.b_vx), r_vx = sh_dn(s.b_vx);
    const double l_vy = sh_up(s.b_vy), r_vy = sh_dn(s.b_vy);
    const double l_p = sh_up(s.b_p), r_p = sh_dn(s.b_p);
    const double l_by = sh_up(s.b_byc), r_by = sh_dn(s.b_byc);
    double xl_rho, xr_rho, xl_vx, xr_vx, xl_vy, xr_vy, xl_p, xr_p, xl_by, xr_by;
    muscl(l_rho, s.b_rho, r_rho, xl_rho, xr_rho);
    muscl(l_vx, s.b_vx, r_vx, xl_vx, xr_vx);
    muscl(l_vy, s.b_vy, r_vy, xl_vy, xr_vy);
    muscl(l_p, s.b_p, r_p, xl_p, xr_p);
    muscl(l_by, s.b_byc, r_by, xl_by, xr_by);
    Flux F;
    {
      Prim L{xr_rho, xr_vx, xr_vy, xr_p, s.bxr_b, xr_by};
      Prim R{sh_dn(xl_rho), sh_dn(xl_vx), sh_dn(xl_vy), sh_dn(xl_p), s.bxr_b, sh_dn(xl_by)};
      F = hll_x(L, R);
    }

    // ---- Ez(k-1) at the bottom-left corner of cell (c, k-1) ----
    const double fby_left = sh_up(F.b);
    const double gbx_left = sh_up(G.b);
    double ez;
    {
      double eN = -fby_left;
      double eS = -s.fby_left;
      double eE = G.b;
      double eW = gbx_left;
      double vxc = 0.25 * (s.b_vx + l_vx + s.a_vx + s.vx_left_a);
      double vyc = 0.25 * (s.b_vy + l_vy + s.a_vy + s.vy_left_a);
      ez = corner_ez(eN, eS, eE, eW, vxc, vyc);
    }

    // ---- stage-2 resistive term: J(c, k-1) of the stage field ----
    double j_b = 0.0;
    if (MODE == 2) {
      const double byl = sh_up(s.byf_b);
      j_b = a.eta * ((s.byf_b - byl) * a.inv_dx - (s.bxl_b - s.bxl_a) * a.inv_dy);
    }

    // ---- dU(k-2): flux divergence, CT update, stage combination ----
    const int j = k - 2;
    const double fl_rho = sh_up(s.f_rho), fl_mx = sh_up(s.f_mx), fl_my = sh_up(s.f_my), fl_E = sh_up(s.f_E);
    const double ez_right = sh_dn(s.ez);
    // the resistive source is a corner quantity: beyond the last corner of an
    // outflow axis it is the last corner's (J(n) = J(n - 1)).  The shuffle
    // runs in every lane before the select: a lane that skipped it would leave
    // its neighbour reading an inactive lane.
    const double j_dn = (MODE == 2) ? sh_dn(s.j_a) : 0.0;
    const double j_right = (MODE == 2) ? ((ox && c0 == nx - 1) ? s.j_a : j_dn) : 0.0;
    const double j_above = (MODE == 2) ? ((oy && k - 2 == ny - 1) ? s.j_a : j_b) : 0.0;
    if (j >= y0 && j < y1 && out_lane) {
      double d[6];
      d[0] = -((s.f_rho - fl_rho) * idx_ + (G.rho - s.g_rho) * idy_);
      d[1] = -((s.f_mx - fl_mx) * idx_ + (G.mx - s.g_mx) * idy_);
      d[2] = -((s.f_my - fl_my) * idx_ + (G.my - s.g_my) * idy_);
      d[3] = -((s.f_E - fl_E) * idx_ + (G.E - s.g_E) * idy_);
      d[4] = -(ez - s.ez) * idy_;
      d[5] = (ez_right - s.ez) * idx_;
      const int o = j * nx + c0;
      if (MODE == 0) {
#pragma unroll
        for (int q = 0; q < 6; ++q) a.o1[q][o] = d[q];
      } else if (MODE == 1) {
        // all point loads before any store: no load has to wait for a store
        double un[6];
        const int old = (slot == 2) ? 0 : slot + 1;    // the slot of row k - 2
#pragma unroll
        for (int q = 0; q < 6; ++q) un[q] = ring[old][q][lane];

#pragma unroll
        for (int q = 0; q < 6; ++q) {
          // W = U^n + dt delta f1 (stage-3 base): next read by stage 2
          __builtin_nontemporal_store(un[q] + d[q] * a.c_b, &a.o1[q][o]);
          double u = un[q] + d[q] * a.c_a;         // U2* = U^n + dt gamma f1
          if (q == 0 && u < kFloor) u = kFloor;
          if (q < 4) __builtin_nontemporal_store(u, &a.o2[q][o]);
          else a.o2[q][o] = u;                     // faces: read next by the solve's start
        }
      } else {
        double wv[6];
#pragma unroll
        for (int q = 0; q < 6; ++q) wv[q] = a.u[q][o];
        double g2x = 0.0, g2y = 0.0;
        if (a.eta != 0.0) {
          g2x = -(j_above - s.j_a) * a.inv_dy;
          g2y = (j_right - s.j_a) * a.inv_dx;
        }
#pragma unroll
        for (int q = 0; q < 6; ++q) {
          double u = wv[q] + d[q] * a.c_b;         // U3* = W + dt (1 - delta) f2
          if (q == 4 && a.eta != 0.0) u = u + g2x * a.c_g;
          if (q == 5 && a.eta != 0.0) u = u + g2y * a.c_g;
          if (q == 0 && u <= kFloor) {
            // the final computePrimitives of the step forms the velocities
            // (which the next CFL step reads) before the floor (Grid.hpp:111-115)
            if (a.rho_raw) a.rho_raw[o] = u;
            if (u < kFloor) u = kFloor;
          }
          if (q < 4) __builtin_nontemporal_store(u, &a.o1[q][o]);
          else a.o1[q][o] = u;                     // faces: read next by the solve's start
        }
      }
    }

    // ---- rotate the carried state ----
    s.a_rho = s.b_rho; s.a_vx = s.b_vx; s.a_vy = s.b_vy; s.a_p = s.b_p; s.a_bxc = s.b_bxc;
    s.b_rho = c_rho; s.b_vx = c_vx; s.b_vy = c_vy; s.b_p = c_p; s.b_bxc = bxc; s.b_byc = byc;
    s.yr_rho = yr_rho; s.yr_vx = yr_vx; s.yr_vy = yr_vy; s.yr_p = yr_p; s.yr_bx = yr_bx;
    s.f_rho = F.rho; s.f_mx = F.mx; s.f_my = F.my; s.f_E = F.E;
    s.fby_left = fby_left;
    s.g_rho = G.rho; s.g_mx = G.mx; s.g_my = G.my; s.g_E = G.E;
    s.ez = ez;
    s.vx_left_a = l_vx; s.vy_left_a = l_vy;
    s.bxr_b = bxr;
    s.byf_b = s.byf_c;
    s.byf_c = byn;
    s.j_a = j_b;
    s.bxl_a = s.bxl_b;
    s.bxl_b = bxl;
    slot = (slot == 2) ? 0 : slot + 1;
  }
}


// The (strip, segment) of this workgroup.  With a.band > 0 the grid is 1D:
// consecutive workgroup ids are dealt to the 8 XCDs in turn, so XCD k runs
// ids k, k + 8, ...; those are renumbered into one contiguous run of virtual
// ids per XCD, and virtual ids walk column bands of a.band strips left to
// right, within a band the rows of strips bottom to top.  Horizontally and
// vertically adjacent strips then share their halo lines in one L2.  Measured
// at 4096^2 (both stages, results bitwise the same): one band (a.band =
// strips: each XCD marches a band of whole rows) cuts the stages' L2 misses
// by 19-22 % and their time by 1-4 % against the 2D grid; narrower bands
// (2, 4, 8, 16 per row) run slower.
__device__ __forceinline__ int2 stage_tile(const StageArgs& a) {
  if (a.band <= 0) return make_int2(blockIdx.x, blockIdx.y);
  const int nb = gridDim.x, b = blockIdx.x;
  const int q = nb >> 3, rem = nb & 7, k = b & 7;
  const int v = k * q + (k < rem ? k : rem) + (b >> 3);
  const int per_band = a.band * a.segs;
  const int bi = v / per_band, r = v - bi * per_band;
  const int left = a.strips - bi * a.band;
  const int w = left < a.band ? left : a.band;
  return make_int2(bi * a.band + r % w, r / w);
}

#ifndef MHD_STAGE_WAVES
#define MHD_STAGE_WAVES 3      // waves per SIMD the single-run stage kernel is compiled for (tuning parameter)
#endif

template <int MODE>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(MHD_STAGE_WAVES))) void k_stage(StageArgs a) {
  const int2 t = stage_tile(a);
  stage_body<MODE>(a, t.x, t.y);
}

template __global__ void k_stage<0>(StageArgs);
template __global__ void k_stage<1>(StageArgs);
template __global__ void k_stage<2>(StageArgs);
// ---------------------------------------------------------------------------
// Reductions.  A kernel that produces a global sum writes one partial per
// block; the kernel that consumes the sum re-reduces all partials itself, in
// a fixed order, in every block.  The kernel boundary is the only
// synchronisation: no device-scope fence or atomic is used (on gfx950 a
// device-scope release writes back the whole L2 of the issuing XCD).  Fixed
// order makes every reduction deterministic.
// ---------------------------------------------------------------------------

template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* lds) {
  // wave-level tree (fixed order), then the wave results in order; result in thread 0
#pragma unroll
  for (int q = 0; q < NV; ++q) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v[q] += __shfl_down(v[q], off, 64);
  }
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) lds[w * NV + q] = v[q];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = blockDim.x >> 6;
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      double t = lds[q];
      for (int ww = 1; ww < nw; ++ww) t += lds[ww * NV + q];
      v[q] = t;
    }
  }
  __syncthreads();
}

__device__ __forceinline__ double max_ignoring_nan(double m, double v) { return (v > m) ? v : m; }

__device__ __forceinline__ void block_max(double& v, double* lds) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = max_ignoring_nan(v, __shfl_down(v, off, 64));
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) lds[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = blockDim.x >> 6;
    double t = lds[0];
    for (int ww = 1; ww < nw; ++ww) t = max_ignoring_nan(t, lds[ww]);
    v = t;
  }
  __syncthreads();
}

// Sum of NV interleaved partial arrays part[b * NV + q] over nb blocks, in a
// fixed order; every thread of the block receives the totals.
template <int NV>
__device__ void sum_partials(const double* __restrict__ part, int nb, double (&out)[NV], double* lds) {
  double v[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) v[q] = 0.0;
  // the loads of kB partials are issued before any is summed (one batch of
  // L2 round trips instead of kB in turn), in the same summation order
  // (4 measured at 4096^2: passes A and B -1 %; 8 the same, 1 the slow end)
#ifndef MHD_PART_BATCH
#define MHD_PART_BATCH 4
#endif
  constexpr int kB = MHD_PART_BATCH;
  for (int b0 = threadIdx.x; b0 < nb; b0 += kB * blockDim.x) {
    double t[kB][NV];
#pragma unroll
    for (int i = 0; i < kB; ++i) {
      const int b = b0 + i * blockDim.x;
#pragma unroll
      for (int q = 0; q < NV; ++q) t[i][q] = (b < nb) ? part[b * NV + q] : 0.0;
    }
#pragma unroll
    for (int i = 0; i < kB; ++i) {
#pragma unroll
      for (int q = 0; q < NV; ++q) v[q] += t[i][q];
    }
  }
  block_sum<NV>(v, lds);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) lds[8 + q] = v[q];
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NV; ++q) out[q] = lds[8 + q];
  __syncthreads();
}

// Tile of a 2D sweep: 256 consecutive columns x `ty` rows.  Tiles are
// numbered with x fastest, so that (with the launch's XCD remap) each XCD
// sweeps a band of whole tile rows and horizontally adjacent tiles, which
// read each other's edge column as their halo, share those lines in its L2.
// Measured at 4096^2 against y-fastest numbering (each XCD two tile columns):
// cg_start -6 %, pass A -4.5 %, pass B -4.3 %, finalize 1 / 2 +1-2 %, the
// step -2.3 %.
struct Tile {
  int vb, i, ic, im, ip, y0, y1;
  bool valid;
};

#ifndef MHD_TILE_XFAST
#define MHD_TILE_XFAST 1      // tile numbering: 1 x fastest, 0 y fastest (tuning parameter)
#endif
// (column, row) of tile vb
__device__ __forceinline__ int2 tile_xy(int nx, int tiles_y, int vb) {
  if (MHD_TILE_XFAST) {
    const int tiles_x = (nx + kTileX - 1) / kTileX;
    const int by = vb / tiles_x;
    return make_int2(vb - by * tiles_x, by);
  }
  const int bx = vb / tiles_y;
  return make_int2(bx, vb - bx * tiles_y);
}

__device__ __forceinline__ Tile tile_of(int nx, int ny, int ty, int tiles_y, int vb, int open_x = 0) {
  Tile t;
  t.vb = vb;
  const int2 xy = tile_xy(nx, tiles_y, vb);
  const int bx = xy.x, by = xy.y;
  t.i = bx * kTileX + threadIdx.x;
  t.valid = t.i < nx;
  t.ic = t.valid ? t.i : nx - 1;
  t.im = bc_wrap(t.ic - 1, nx, open_x);
  t.ip = bc_wrap(t.ic + 1, nx, open_x);
  t.y0 = by * ty;
  t.y1 = min(t.y0 + ty, ny);
  return t;
}

// ---------------------------------------------------------------------------
// Matrix-free Jacobi-preconditioned CG on the corner unknowns, restating
// Eigen's ConjugateGradient (zero initial guess, stop when
// |r|^2 < max(tol^2 |b|^2, DBL_MIN), iteration count = completed non-final
// iterations, maxIter = 2 N) on the operator of generate_A_matrix
// (solvers.hpp:404-443):
//   (A x)_ij = d x_ij - cx (x_{i-1,j} + x_{i+1,j}) - cy (x_{i,j-1} + x_{i,j+1}),
//   d = 1 + 2 (cx + cy), cx = alpha / dx^2, cy = alpha / dy^2, periodic.
// The Jacobi preconditioner is the constant 1/d.  Kernel sequence of a solve:
//   init, A(0), B(0), A(1), B(1), ...
// A(it) reduces the partials of the previous kernel, takes the convergence
// decision and beta, then forms p = invd r + beta p_old and sums p . Ap.
// B(it) forms alpha, x += alpha p, r -= alpha Ap and sums |r|^2, r . (invd r).
// The stencil product is recomputed in B instead of being stored.  Block 0
// records the scalar state; scalars written by a kernel are read only by
// later kernels (abs_new is double-buffered by iteration parity).
// ---------------------------------------------------------------------------

// A p with the accumulation order of a row of the assembled operator
// (south, west, centre, east, north), identical in every pass.
__device__ __forceinline__ double stencil(double ps, double pw, double pc, double pe, double pn, double cx,
                                          double cy, double d) {
  double q = -cy * ps;
  q = fma(-cx, pw, q);
  q = fma(d, pc, q);
  q = fma(-cx, pe, q);
  q = fma(-cy, pn, q);
  return q;
}

// Columns of a tile: `w` valid columns starting at x0 and the periodic left
// and right halo columns xl, xr.
__device__ __forceinline__ void tile_columns(int nx, int tiles_y, int vb, int& x0, int& w, int& xl, int& xr,
                                             int open_x = 0) {
  const int bx = tile_xy(nx, tiles_y, vb).x;
  x0 = bx * kTileX;
  w = min(kTileX, nx - x0);
  xl = bc_wrap(x0 - 1, nx, open_x);
  xr = bc_wrap(x0 + w, nx, open_x);
}

// eta J(B) at corner (i, j) (resistive_B_term's first loop, solvers.hpp:450-463)
// from the face values by(i, j), by(i-1, j), bx(i, j), bx(i, j-1).
__device__ __forceinline__ double rhs_from(const CgArgs& a, double by_c, double by_w, double bx_c, double bx_s) {
  const double dBy_dx = (by_c - by_w) * a.inv_dx;
  const double dBx_dy = (bx_c - bx_s) * a.inv_dy;
  return a.eta * (dBy_dx - dBx_dy);
}
template <bool OPEN = true>
__device__ __forceinline__ double rhs_at(const CgArgs& a, int i, int j) {
  const int nx = a.nx, ny = a.ny;
  const int im = bc_wrap(i - 1, nx, OPEN ? a.open_x : 0);
  const int jm = bc_wrap(j - 1, ny, OPEN ? a.open_y : 0);
  return rhs_from(a, a.by[j * nx + i], a.by[j * nx + im], a.bx[j * nx + i], a.bx[jm * nx + i]);
}

__device__ __forceinline__ int wrap_row(int j, int ny, int open_y = 0) { return bc_wrap(j, ny, open_y); }

// Jacobi weight 1 / A_ii of corner (i, j).  On an outflow axis a boundary
// corner's clamped neighbour folds into the diagonal (A_ii = d - c), summed in
// the order the assembled operator's entries of that row appear (south,
// north, west, east, then the diagonal).
__device__ __forceinline__ double invd_at(const CgArgs& a, int i, int j) {
  if (!(a.open_x | a.open_y)) return a.invd;
  double s = 0.0;
  bool edge = false;
  if (a.open_y && j == 0) { s += -a.cy; edge = true; }
  if (a.open_y && j == a.ny - 1) { s += -a.cy; edge = true; }
  if (a.open_x && i == 0) { s += -a.cx; edge = true; }
  if (a.open_x && i == a.nx - 1) { s += -a.cx; edge = true; }
  return edge ? 1.0 / (s + a.diag) : a.invd;
}

// The bodies of the CG passes are instantiated per boundary mode: with both
// axes periodic the Jacobi weight is the constant 1 / d and nothing of the
// outflow handling is in the loop.
template <bool OPEN>
__device__ __forceinline__ double invd_of(const CgArgs& a, int i, int j) {
  if (!OPEN) return a.invd;
  return invd_at(a, i, j);
}

// The 2D sweeps below march a tile one row per iteration (pass B downwards,
// the others upwards: all tiles run at once, so consecutive passes meet the
// rows the previous pass touched last while the memory-side cache still holds
// them).  Global loads are issued two rows ahead of their use, so each wave
// keeps a row of loads in flight across the LDS exchange and the arithmetic
// of the current row (the barrier only waits for LDS traffic).  Data that is
// not read again before it would be evicted (x in pass B, p_old in pass A)
// moves with nontemporal loads and stores.

// Start of a solve: b = eta J(B) (or bin), r = b, p = invd b, and the first
// product q = A p; partials |b|^2, b . p, p . q.
template <bool FROM_B, bool OPEN>
__device__ __forceinline__ void cg_start_body(const CgArgs& a, const double* __restrict__ bin, int vb) {
  __shared__ double prow[2][kTileX + 2];
  __shared__ double lds[16];
  const int nx = a.nx, ny = a.ny, oy = (OPEN ? a.open_y : 0);
  const Tile t = tile_of(nx, ny, a.ty, a.tiles_y, vb, (OPEN ? a.open_x : 0));
  int x0, w, xl, xr;
  tile_columns(nx, a.tiles_y, vb, x0, w, xl, xr, (OPEN ? a.open_x : 0));
  const int tid = threadIdx.x;
  const bool left = (tid == 0), right = (tid == w - 1);
  const double d = a.diag, cx = a.cx, cy = a.cy;
  // b of row j for the own column, with bx of the row below carried (rows are
  // visited in order; on an outflow axis the corner row below the first is the
  // first, that above the last is the last)
  double bx_carry = FROM_B ? a.bx[wrap_row(t.y0 - 2, ny, oy) * nx + t.ic] : 0.0;
  double b_last = 0.0;
  auto own_b = [&](int j) -> double {
    const int r = wrap_row(j, ny, oy);
    if (!FROM_B) return bin[r * nx + t.ic];
    if (oy && j >= ny) return b_last;
    const double bxc = a.bx[r * nx + t.ic];
    const double v = rhs_from(a, a.by[r * nx + t.ic], a.by[r * nx + t.im], bxc, bx_carry);
    bx_carry = bxc;
    b_last = v;
    return v;
  };
  // b of a halo column (rows visited in order from the tile's first), with bx
  // of the row below carried in cb as for the own column
  auto halo_b = [&](int col, int j, double& cb) -> double {
    const int r = wrap_row(j, ny, oy);
    if (!FROM_B) return bin[r * nx + col];
    const int cm = bc_wrap(col - 1, nx, (OPEN ? a.open_x : 0));
    const double bxc = a.bx[r * nx + col];
    const double v = rhs_from(a, a.by[r * nx + col], a.by[r * nx + cm], bxc, cb);
    cb = bxc;
    return v;
  };
  double cbl = (FROM_B && left) ? a.bx[wrap_row(t.y0 - 1, ny, oy) * nx + xl] : 0.0;
  double cbr = (FROM_B && right) ? a.bx[wrap_row(t.y0 - 1, ny, oy) * nx + xr] : 0.0;
  double acc[3] = {0.0, 0.0, 0.0};
  double bm = own_b(t.y0 - 1);
  double bc = own_b(t.y0);
  double bn = own_b(t.y0 + 1);
  double hl = left ? halo_b(xl, t.y0, cbl) : 0.0;
  double hr = right ? halo_b(xr, t.y0, cbr) : 0.0;
  for (int j = t.y0; j < t.y1; ++j) {
    // loads for the next row
    const double bnn = (j + 1 < t.y1) ? own_b(j + 2) : 0.0;
    const double hl_n = (left && j + 1 < t.y1) ? halo_b(xl, j + 1, cbl) : 0.0;
    const double hr_n = (right && j + 1 < t.y1) ? halo_b(xr, j + 1, cbr) : 0.0;
    const int buf = (j - t.y0) & 1;
    const double pc = invd_of<OPEN>(a, t.ic, j) * bc;
    if (t.valid) prow[buf][tid + 1] = pc;
    if (left) prow[buf][0] = invd_of<OPEN>(a, xl, j) * hl;
    if (right) prow[buf][w + 1] = invd_of<OPEN>(a, xr, j) * hr;
    __syncthreads();
    const double pm = invd_of<OPEN>(a, t.ic, wrap_row(j - 1, ny, oy)) * bm;
    const double pn = invd_of<OPEN>(a, t.ic, wrap_row(j + 1, ny, oy)) * bn;
    const double q = stencil(pm, prow[buf][tid], pc, prow[buf][tid + 2], pn, cx, cy, d);
    if (t.valid) {
      const int o = j * nx + t.ic;
      a.r[o] = bc - 0.0;
      a.p0[o] = pc;
      acc[0] += bc * bc;
      acc[1] += bc * pc;
      acc[2] += pc * q;
    }
    bm = bc;
    bc = bn;
    bn = bnn;
    hl = hl_n;
    hr = hr_n;
  }
  block_sum<3>(acc, lds);
  if (tid == 0) {
    a.part_s[t.vb * 3 + 0] = acc[0];
    a.part_s[t.vb * 3 + 1] = acc[1];
    a.part_s[t.vb * 3 + 2] = acc[2];
  }
}

// Right-hand side only (explicit resistive term): r = eta J(B).
__global__ __launch_bounds__(kTileX) void k_cg_rhs(CgArgs a) {
  const int nx = a.nx, ny = a.ny;
  const Tile t = tile_of(nx, ny, a.ty, a.tiles_y, xcd_block(blockIdx.x, gridDim.x), a.open_x);
  if (!t.valid) return;
  for (int j = t.y0; j < t.y1; ++j) a.r[j * nx + t.ic] = rhs_at(a, t.ic, j) - 0.0;
}

// Pass A(it), it >= 1: convergence decision on the partials of B(it - 1),
// beta, p_new = invd r + beta p_old, q = A p_new; partial p_new . q.
template <bool OPEN>
__device__ __forceinline__ void cg_pass_a_body(const CgArgs& a, const double* __restrict__ pold,
                                               double* __restrict__ pnew, int it, int vb, int nb) {
  __shared__ double prow[2][kTileX + 2];
  __shared__ double lds[16];
  CgState* st = a.st;
  if (st->done) return;
  double rr[2];
  sum_partials<2>(a.part_b, nb, rr, lds);
  const Tile t = tile_of(a.nx, a.ny, a.ty, a.tiles_y, vb, (OPEN ? a.open_x : 0));
  const bool lead = (t.vb == 0) && (threadIdx.x == 0);
  const double rn2 = rr[0];
  if (rn2 < st->threshold) {                // converged after B(it - 1)
    if (lead) { st->rn2 = rn2; st->iters = it - 1; st->done = 1; }
    return;
  }
  const double beta = rr[1] / st->abs_new[(it - 1) & 1];
  if (lead) { st->rn2 = rn2; st->iters = it; st->abs_new[it & 1] = rr[1]; st->beta = beta; }
  if (it >= st->max_iters) {
    if (lead) st->done = 2;
    return;
  }
  const int nx = a.nx, ny = a.ny, oy = (OPEN ? a.open_y : 0);
  int x0, w, xl, xr;
  tile_columns(nx, a.tiles_y, vb, x0, w, xl, xr, (OPEN ? a.open_x : 0));
  const int tid = threadIdx.x;
  const bool left = (tid == 0), right = (tid == w - 1);
  const double d = a.diag, cx = a.cx, cy = a.cy;
  auto pv = [&](double po, double ro, int col, int j) -> double {
    return fma(beta, po, invd_of<OPEN>(a, col, wrap_row(j, ny, oy)) * ro);
  };
  auto at = [&](int col, int j) -> int { return wrap_row(j, ny, oy) * nx + col; };
  double acc[1] = {0.0};
  double pm = pv(__builtin_nontemporal_load(&pold[at(t.ic, t.y0 - 1)]), a.r[at(t.ic, t.y0 - 1)], t.ic, t.y0 - 1);
  double pc = pv(__builtin_nontemporal_load(&pold[at(t.ic, t.y0)]), a.r[at(t.ic, t.y0)], t.ic, t.y0);
  double n_po = __builtin_nontemporal_load(&pold[at(t.ic, t.y0 + 1)]), n_ro = a.r[at(t.ic, t.y0 + 1)];
  double l_po = left ? __builtin_nontemporal_load(&pold[at(xl, t.y0)]) : 0.0, l_ro = left ? a.r[at(xl, t.y0)] : 0.0;
  double r_po = right ? __builtin_nontemporal_load(&pold[at(xr, t.y0)]) : 0.0, r_ro = right ? a.r[at(xr, t.y0)] : 0.0;
  for (int j = t.y0; j < t.y1; ++j) {
    const bool more = j + 1 < t.y1;
    // loads for the next row
    // own column unconditionally (last row: a valid wrapped row, unused; 1 %
    // faster than the conditional form, which leaves a branch between the
    // row's memory accesses -- the same change slows pass B and the start
    // kernel by 4-5 %)
    const double nn_po = __builtin_nontemporal_load(&pold[at(t.ic, j + 2)]);
    const double nn_ro = a.r[at(t.ic, j + 2)];
    const double ln_po = (left && more) ? __builtin_nontemporal_load(&pold[at(xl, j + 1)]) : 0.0;
    const double ln_ro = (left && more) ? a.r[at(xl, j + 1)] : 0.0;
    const double rn_po = (right && more) ? __builtin_nontemporal_load(&pold[at(xr, j + 1)]) : 0.0;
    const double rn_ro = (right && more) ? a.r[at(xr, j + 1)] : 0.0;
    const double pn = pv(n_po, n_ro, t.ic, j + 1);
    const int buf = (j - t.y0) & 1;
    if (t.valid) prow[buf][tid + 1] = pc;
    if (left) prow[buf][0] = pv(l_po, l_ro, xl, j);
    if (right) prow[buf][w + 1] = pv(r_po, r_ro, xr, j);
    __syncthreads();
    const double q = stencil(pm, prow[buf][tid], pc, prow[buf][tid + 2], pn, cx, cy, d);
    if (t.valid) {
      pnew[j * nx + t.ic] = pc;
      acc[0] += pc * q;
    }
    pm = pc;
    pc = pn;
    n_po = nn_po; n_ro = nn_ro;
    l_po = ln_po; l_ro = ln_ro;
    r_po = rn_po; r_ro = rn_ro;
  }
  block_sum<1>(acc, lds);
  if (tid == 0) a.part_a[t.vb] = acc[0];
}

// Pass B(it): alpha = absNew / (p . Ap), x += alpha p, r -= alpha Ap;
// partials |r|^2 and r . (invd r).  B(0) also takes the start-of-solve
// decisions of Eigen's conjugate_gradient from the partials of the start kernel.
template <bool FIRST, bool OPEN>
__device__ __forceinline__ void cg_pass_b_body(const CgArgs& a, const double* __restrict__ pv, int it, int vb, int nb) {
  __shared__ double prow[2][kTileX + 2];
  __shared__ double lds[16];
  CgState* st = a.st;
  if (st->done) return;
  const Tile t = tile_of(a.nx, a.ny, a.ty, a.tiles_y, vb, (OPEN ? a.open_x : 0));
  const bool lead = (t.vb == 0) && (threadIdx.x == 0);
  double alpha, pap;
  if (FIRST) {
    double s3[3];
    sum_partials<3>(a.part_s, nb, s3, lds);
    const double rhs = s3[0];
    const double thr = dmax(a.tol * a.tol * rhs, 2.2250738585072014e-308);
    int done = 0;
    if (rhs == 0.0) done = 3;                 // x = 0, no iteration
    else if (rhs < thr) done = 1;             // residual below threshold at x = 0
    else if (st->max_iters <= 0) done = 2;
    if (lead) {
      st->rhs_norm2 = rhs;
      st->threshold = thr;
      st->rn2 = rhs;
      st->abs_new[0] = s3[1];
      st->iters = 0;
      if (done) st->done = done;
    }
    if (done) return;
    pap = s3[2];
    alpha = s3[1] / pap;
  } else {
    double s1[1];
    sum_partials<1>(a.part_a, nb, s1, lds);
    pap = s1[0];
    alpha = st->abs_new[it & 1] / pap;
  }
  if (lead) {
    st->pap = pap;
    st->alpha = alpha;
    st->spmv = it + 1;
  }
  const int nx = a.nx, ny = a.ny;
  int x0, w, xl, xr;
  tile_columns(nx, a.tiles_y, vb, x0, w, xl, xr, (OPEN ? a.open_x : 0));
  const int tid = threadIdx.x;
  const bool left = (tid == 0), right = (tid == w - 1);
  const double d = a.diag, cx = a.cx, cy = a.cy;
  auto at = [&](int col, int j) -> int { return wrap_row(j, ny, (OPEN ? a.open_y : 0)) * nx + col; };
  double acc[2] = {0.0, 0.0};
  // rows are swept downwards (pass A sweeps upwards), so that this pass first
  // reads the rows of p and r that the previous pass wrote last
  const int j0 = t.y1 - 1, nrow = t.y1 - t.y0;
  double pa = pv[at(t.ic, j0 + 1)];           // north of the current row
  double pc = pv[at(t.ic, j0)];
  double pb = pv[at(t.ic, j0 - 1)];           // south
  double xc = FIRST ? 0.0 : __builtin_nontemporal_load(&a.x[j0 * nx + t.ic]);
  double rc = a.r[j0 * nx + t.ic];
  double hl = left ? pv[at(xl, j0)] : 0.0;
  double hr = right ? pv[at(xr, j0)] : 0.0;
  for (int k = 0; k < nrow; ++k) {
    const int j = j0 - k;
    const bool more = k + 1 < nrow;
    // loads for the next row (j - 1)
    const double pbb = more ? pv[at(t.ic, j - 2)] : 0.0;
    const double xn = (!FIRST && more) ? __builtin_nontemporal_load(&a.x[(j - 1) * nx + t.ic]) : 0.0;
    const double rnx = more ? a.r[(j - 1) * nx + t.ic] : 0.0;
    const double hl_n = (left && more) ? pv[at(xl, j - 1)] : 0.0;
    const double hr_n = (right && more) ? pv[at(xr, j - 1)] : 0.0;
    const int buf = k & 1;
    if (t.valid) prow[buf][tid + 1] = pc;
    if (left) prow[buf][0] = hl;
    if (right) prow[buf][w + 1] = hr;
    __syncthreads();
    const double q = stencil(pb, prow[buf][tid], pc, prow[buf][tid + 2], pa, cx, cy, d);
    if (t.valid) {
      const int o = j * nx + t.ic;
      __builtin_nontemporal_store(xc + alpha * pc, &a.x[o]);
      const double rn = rc - alpha * q;
      a.r[o] = rn;
      const double z = invd_of<OPEN>(a, t.ic, j) * rn;
      acc[0] += rn * rn;
      acc[1] += rn * z;
    }
    pa = pc;
    pc = pb;
    pb = pbb;
    xc = xn;
    rc = rnx;
    hl = hl_n;
    hr = hr_n;
  }
  block_sum<2>(acc, lds);
  if (tid == 0) {
    a.part_b[t.vb * 2 + 0] = acc[0];
    a.part_b[t.vb * 2 + 1] = acc[1];
  }
}

// Periodic kernels, and separate kernels for contexts with an outflow axis
// (one body per kernel: a kernel holding both would be compiled for the
// heavier one, which costs the periodic passes a quarter of their speed).
template <bool FROM_B>
__global__ __launch_bounds__(kTileX) void k_cg_start(CgArgs a, const double* __restrict__ bin) {
  cg_start_body<FROM_B, false>(a, bin, xcd_block(blockIdx.x, gridDim.x));
}
__global__ __launch_bounds__(kTileX) void k_cg_pass_a(CgArgs a, const double* __restrict__ pold,
                                                      double* __restrict__ pnew, int it) {
  cg_pass_a_body<false>(a, pold, pnew, it, xcd_block(blockIdx.x, gridDim.x), gridDim.x);
}
template <bool FIRST>
__global__ __launch_bounds__(kTileX) void k_cg_pass_b(CgArgs a, const double* __restrict__ pv, int it) {
  cg_pass_b_body<FIRST, false>(a, pv, it, xcd_block(blockIdx.x, gridDim.x), gridDim.x);
}
template <bool FROM_B>
__global__ __launch_bounds__(kTileX) void k_cg_start_open(CgArgs a, const double* __restrict__ bin) {
  cg_start_body<FROM_B, true>(a, bin, xcd_block(blockIdx.x, gridDim.x));
}
__global__ __launch_bounds__(kTileX) void k_cg_pass_a_open(CgArgs a, const double* __restrict__ pold,
                                                           double* __restrict__ pnew, int it) {
  cg_pass_a_body<true>(a, pold, pnew, it, xcd_block(blockIdx.x, gridDim.x), gridDim.x);
}
template <bool FIRST>
__global__ __launch_bounds__(kTileX) void k_cg_pass_b_open(CgArgs a, const double* __restrict__ pv, int it) {
  cg_pass_b_body<FIRST, true>(a, pv, it, xcd_block(blockIdx.x, gridDim.x), gridDim.x);
}
template __global__ void k_cg_start_open<true>(CgArgs, const double*);
template __global__ void k_cg_start_open<false>(CgArgs, const double*);
template __global__ void k_cg_pass_b_open<true>(CgArgs, const double*, int);
template __global__ void k_cg_pass_b_open<false>(CgArgs, const double*, int);

template __global__ void k_cg_start<true>(CgArgs, const double*);
template __global__ void k_cg_start<false>(CgArgs, const double*);
template __global__ void k_cg_pass_b<true>(CgArgs, const double*, int);
template __global__ void k_cg_pass_b<false>(CgArgs, const double*, int);

// ---------------------------------------------------------------------------
// Implicit-stage finalisers.
// ---------------------------------------------------------------------------

// rate = -curl_z(x): rate_x(i,j) = -(x(i,j+1) - x(i,j))/dy, rate_y = (x(i+1,j) - x(i,j))/dx
// (resistive_B_term's second loop, solvers.hpp:472-488).

// B += rate(x) * c (solvers.hpp:528-529).  x is taken as 0 when CG did no
// product (Eigen returns x = 0), which leaves B unchanged.
__device__ __forceinline__ void finalize1_body(const FinArgs& a, int vb) {
  const int nx = a.nx, ny = a.ny;
  if (a.st->spmv == 0) {            // x = 0: B unchanged (copied when the output is another array)
    if (a.bx_out == a.bx) return;
    const Tile t = tile_of(nx, ny, a.ty, a.tiles_y, vb);
    if (!t.valid) return;
    for (int j = t.y0; j < t.y1; ++j) {
      const int o = j * nx + t.ic;
      a.bx_out[o] = a.bx[o];
      a.by_out[o] = a.by[o];
    }
    return;
  }
  const Tile t = tile_of(nx, ny, a.ty, a.tiles_y, vb, a.open_x);
  if (!t.valid) return;
  auto at = [&](int col, int j) -> int { return wrap_row(j, ny, a.open_y) * nx + col; };
  double xc = a.x[at(t.ic, t.y0)], xn = a.x[at(t.ic, t.y0 + 1)], xe = a.x[at(t.ip, t.y0)];
  double bxv = a.bx[t.y0 * nx + t.ic], byv = a.by[t.y0 * nx + t.ic];
  for (int j = t.y0; j < t.y1; ++j) {
    const bool more = j + 1 < t.y1;
    // loads for the next row (rows j + 1 of B are read before row j is written)
    const double xnn = more ? a.x[at(t.ic, j + 2)] : 0.0;
    const double xen = more ? a.x[at(t.ip, j + 1)] : 0.0;
    const double bxn = more ? a.bx[(j + 1) * nx + t.ic] : 0.0;
    const double byn = more ? a.by[(j + 1) * nx + t.ic] : 0.0;
    const int o = j * nx + t.ic;
    a.bx_out[o] = bxv + (-(xn - xc) * a.inv_dy) * a.c;
    a.by_out[o] = byv + ((xe - xc) * a.inv_dx) * a.c;
    xc = xn; xn = xnn; xe = xen; bxv = bxn; byv = byn;
  }
}

// Final recombination of the B field (solvers.hpp:568-608): with
// B3 = B3* + c rate(x), B^{n+1} = B3* + c g3, where g3 = -curl(eta J(B3)) is
// the explicit resistive term of B3 (solvers.hpp:579) -- the reference
// re-evaluates g3 rather than copying B3.  Each thread marches its column up
// the tile computing J3 = eta J(B3) at the corners of the next row; J3 of the
// column to the right comes from an LDS row written one step earlier (the
// thread at the right edge of the tile also runs the halo column).
struct Fin2Chain {
  double bx3_below;   // bx3(col, row - 1)
  double bxs, bys;    // B3* faces of the last row computed
};

// The loads of J3(col, row): x at (col, row), (col, row + 1), (col + 1, row),
// (col - 1, row) and the B3* faces bx(col, row), by(col, row), by(col - 1, row).
struct Fin2Loads {
  double xc, xn, xe, xw, bxs, bys, bysw;
};

template <bool OPEN>
__device__ __forceinline__ Fin2Loads fin2_load(const FinArgs& a, bool xv, int col, int row) {
  const int nx = a.nx, ny = a.ny;
  const int cm = bc_wrap(col - 1, nx, (OPEN ? a.open_x : 0));
  const int cp = bc_wrap(col + 1, nx, (OPEN ? a.open_x : 0));
  const int rp = bc_wrap(row + 1, ny, (OPEN ? a.open_y : 0));
  Fin2Loads v;
  v.xc = xv ? a.x[row * nx + col] : 0.0;
  v.xn = xv ? a.x[rp * nx + col] : 0.0;
  v.xe = xv ? a.x[row * nx + cp] : 0.0;
  v.xw = xv ? a.x[row * nx + cm] : 0.0;
  v.bxs = a.bx[row * nx + col];
  v.bys = a.by[row * nx + col];
  v.bysw = a.by[row * nx + cm];
  return v;
}

// The loads of the row after the one prev was loaded for: x there is prev's
// x of the row above (the row sequence of finalize2_body advances by one,
// wrapped or clamped as bc_wrap does), so it is not loaded again.
template <bool OPEN>
__device__ __forceinline__ Fin2Loads fin2_load_after(const FinArgs& a, bool xv, int col, int row,
                                                     const Fin2Loads& prev) {
  const int nx = a.nx, ny = a.ny;
  const int cm = bc_wrap(col - 1, nx, (OPEN ? a.open_x : 0));
  const int cp = bc_wrap(col + 1, nx, (OPEN ? a.open_x : 0));
  const int rp = bc_wrap(row + 1, ny, (OPEN ? a.open_y : 0));
  Fin2Loads v;
  v.xc = prev.xn;
  v.xn = xv ? a.x[rp * nx + col] : 0.0;
  v.xe = xv ? a.x[row * nx + cp] : 0.0;
  v.xw = xv ? a.x[row * nx + cm] : 0.0;
  v.bxs = a.bx[row * nx + col];
  v.bys = a.by[row * nx + col];
  v.bysw = a.by[row * nx + cm];
  return v;
}

// J3(col, row) from its loads, given chain.bx3_below = bx3(col, row - 1);
// advances the chain.
template <bool OPEN>
__device__ __forceinline__ double fin2_j3_of(const FinArgs& a, int col, const Fin2Loads& v, Fin2Chain& ch) {
  const double bx3 = v.bxs + (-(v.xn - v.xc) * a.inv_dy) * a.c;     // bx3(col, row)
  const double by3 = v.bys + ((v.xe - v.xc) * a.inv_dx) * a.c;      // by3(col, row)
  // by3(col - 1, row); left of the first column of an outflow axis B3 is
  // extended by its boundary value (the ghost rule on the recombined field)
  const double by3w = ((OPEN ? a.open_x : 0) && col == 0) ? by3 : v.bysw + ((v.xc - v.xw) * a.inv_dx) * a.c;
  const double dBy_dx = (by3 - by3w) * a.inv_dx;
  const double dBx_dy = (bx3 - ch.bx3_below) * a.inv_dy;
  ch.bx3_below = bx3;
  ch.bxs = v.bxs;
  ch.bys = v.bys;
  return a.eta * (dBy_dx - dBx_dy);
}

template <bool OPEN>
__device__ __forceinline__ double fin2_j3(const FinArgs& a, bool xv, int col, int row, Fin2Chain& ch) {
  return fin2_j3_of<OPEN>(a, col, fin2_load<OPEN>(a, xv, col, row), ch);
}

template <bool OPEN>
__device__ __forceinline__ Fin2Chain fin2_begin(const FinArgs& a, bool xv, int col, int row0) {
  const int nx = a.nx, ny = a.ny;
  // bx3 of the row below row0; below the first row of an outflow axis, that
  // of the first row (the ghost rule on the recombined field)
  const int rm = ((OPEN ? a.open_y : 0) && row0 == 0) ? 0 : wrap1(row0 - 1, ny);
  const int rmn = bc_wrap(rm + 1, ny, (OPEN ? a.open_y : 0));
  const double xc = xv ? a.x[rm * nx + col] : 0.0;
  const double xn = xv ? a.x[rmn * nx + col] : 0.0;
  Fin2Chain ch;
  ch.bx3_below = a.bx[rm * nx + col] + (-(xn - xc) * a.inv_dy) * a.c;
  ch.bxs = ch.bys = 0.0;
  return ch;
}

// CFL_condition_checker's wave speed of one cell (diagnostics.hpp:58-80) from
// its conserved values and cell-centred B: max into m, non-finite count into bad.
__device__ __forceinline__ void cfl_cell(double rho, double mx, double my, double E, double bxc, double byc,
                                         double& m, double& bad) {
  double rf, vx, vy, p;
  primitives(rho, mx, my, E, bxc, byc, rf, vx, vy, p);
  rf = dmax(rf, 1e-8);
  p = dmax(p, 1e-8);
  // the two divisions by rho as one refined reciprocal, the square roots by
  // sqrt_nr (each within about an ulp; rf >= 1e-8 and ss + as >= 1.6e-16 are
  // normal, a speed below 2^-500 adds nothing to them); non-finite inputs give
  // a non-finite speed as the IEEE forms do
  const double inv = rcp(rf);
  const double ss = (kGamma * p) * inv;
  const double as = (bxc * bxc + byc * byc) * inv;
  const double v2 = vx * vx + vy * vy;
  const double cur = sqrt_nr(ss + as) + ((v2 >= 0x1p-1000) ? sqrt_nr(v2) : 0.0);
  if (cur > m) m = cur;
  if (!(fabs(cur) <= 1.7976931348623157e308)) bad += 1.0;
}

// CFL == true: the tile also produces the CFL partials of the new state (the
// per-tile max speed and non-finite count of cfl_body, from the B it has just
// formed and d.s[0..3] = rho, mx, my, E of the new state) into d.part.  The
// march then runs one row further (the new by of row y1 closes the top row's
// cell-centred by) and the CFL of row j - 1 is formed at row j; the new bx of
// the column to the right comes through LDS, that of the halo column from the
// edge thread's own chain.
template <bool CFL, bool OPEN>
__device__ __forceinline__ void finalize2_body(const FinArgs& a, int vb, const DiagArgs& d) {
  __shared__ double jrow[2][kTileX + 2];
  __shared__ double bxrow[CFL ? 2 : 1][CFL ? kTileX + 2 : 1];
  __shared__ double lds[16];
  const bool xv = a.st->spmv != 0;
  const int nx = a.nx, ny = a.ny;
  const Tile t = tile_of(nx, ny, a.ty, a.tiles_y, vb, (OPEN ? a.open_x : 0));
  int x0, w, xl, xr;
  tile_columns(nx, a.tiles_y, vb, x0, w, xl, xr, (OPEN ? a.open_x : 0));
  (void)xl;
  const int tid = threadIdx.x;
  const bool edge = (tid == w - 1);
  Fin2Chain own = fin2_begin<OPEN>(a, xv, t.ic, t.y0);
  Fin2Chain hal = fin2_begin<OPEN>(a, xv, xr, t.y0);
  double jc = fin2_j3<OPEN>(a, xv, t.ic, t.y0, own);
  double jhc = 0.0;                              // J3(xr, j) of the halo chain (CFL only)
  if (t.valid) jrow[0][tid + 1] = jc;
  if (edge) {
    const double jh = fin2_j3<OPEN>(a, xv, xr, t.y0, hal);
    jrow[0][w + 1] = jh;
    jhc = jh;
  }
  __syncthreads();
  double m = 1e-8, bad = 0.0;
  double bxo_prev = 0.0, byo_prev = 0.0;
  double u_prev[4] = {0.0, 0.0, 0.0, 0.0};     // rho, mx, my, E of row j - 1 (loaded one row ahead)
  const int y_end = CFL ? t.y1 + 1 : t.y1;
  // row j wrapped (the extra row of a top tile; the last row on an outflow axis)
  auto wrapped = [&](int j) -> int { return (j >= ny) ? ((OPEN ? a.open_y : 0) ? ny - 1 : j - ny) : j; };
  auto above = [&](int j) -> int { return bc_wrap(wrapped(j) + 1, ny, (OPEN ? a.open_y : 0)); };
  // the loads of J3 one row ahead of their use
  Fin2Loads nxt = fin2_load<OPEN>(a, xv, t.ic, above(t.y0));
  Fin2Loads hnxt{};
  if (edge) hnxt = fin2_load<OPEN>(a, xv, xr, above(t.y0));
  for (int j = t.y0; j < y_end; ++j) {
    double u_row[4] = {0.0, 0.0, 0.0, 0.0};
    if (CFL) {
      // unconditional (the extra row past the tile reloads its last row; its
      // values are not used), so that no branch separates these loads
      const int o = (j < t.y1 ? j : t.y1 - 1) * nx + t.ic;
#pragma unroll
      for (int q = 0; q < 4; ++q) u_row[q] = d.s[q][o];
    }
    const Fin2Loads cur = nxt, hcur = hnxt;
    if (j + 1 < y_end) {
      nxt = fin2_load_after<OPEN>(a, xv, t.ic, above(j + 1), cur);
      if (edge) hnxt = fin2_load_after<OPEN>(a, xv, xr, above(j + 1), hcur);
    }
    const int buf = (j - t.y0) & 1;
    const int jw = wrapped(j);
    // J3 above the last corner row of an outflow axis is the last row's
    const bool top = (OPEN ? a.open_y : 0) && jw == ny - 1;
    const double bxs = own.bxs, bys = own.bys;   // B3* faces of row j
    const double hbxs = hal.bxs;                 // B3* bx of row j, halo column (edge thread)
    const double jn = fin2_j3_of<OPEN>(a, t.ic, cur, own);
    if (t.valid) jrow[buf ^ 1][tid + 1] = jn;
    double jhn = 0.0;
    if (edge) {
      jhn = fin2_j3_of<OPEN>(a, xr, hcur, hal);
      jrow[buf ^ 1][w + 1] = jhn;
    }
    const double jr = jrow[buf][tid + 2];        // J3(i + 1, j)
    const double g3x = -((top ? jc : jn) - jc) * a.inv_dy;
    const double g3y = (jr - jc) * a.inv_dx;
    const double bxo = bxs + g3x * a.c;
    const double byo = bys + g3y * a.c;
    if (t.valid && j < t.y1) {
      const int o = j * nx + t.ic;
      a.bx_out[o] = bxo;
      a.by_out[o] = byo;
    }
    if (CFL) {
      if (t.valid) bxrow[buf][tid + 1] = bxo;
      if (edge) bxrow[buf][w + 1] = hbxs + (-((top ? jhc : jhn) - jhc) * a.inv_dy) * a.c;   // new bx(xr, j)
      if (j > t.y0 && t.valid) {
        // cell (i, j - 1): bx faces (i, j - 1), (i + 1, j - 1); by faces (i, j - 1), (i, j);
        // above the last row of an outflow axis by is the last row's
        const double bxc = (bxo_prev + bxrow[buf ^ 1][tid + 2]) / 2.0;
        const double byc = (byo_prev + (((OPEN ? a.open_y : 0) && j >= ny) ? byo_prev : byo)) / 2.0;
        const double rr = (d.rho_raw && u_prev[0] == kFloor) ? d.rho_raw[(j - 1) * nx + t.ic] : u_prev[0];
        cfl_cell(rr, u_prev[1], u_prev[2], u_prev[3], bxc, byc, m, bad);
      }
      bxo_prev = bxo;
      byo_prev = byo;
      jhc = jhn;
#pragma unroll
      for (int q = 0; q < 4; ++q) u_prev[q] = u_row[q];
    }
    __syncthreads();
    jc = jn;
  }
  if (CFL) {
    double b1[1] = {bad};
    block_max(m, lds);
    block_sum<1>(b1, lds + 8);
    if (threadIdx.x == 0) {
      d.part[t.vb * 2 + 0] = m;
      d.part[t.vb * 2 + 1] = b1[0];
    }
  }
}

__global__ __launch_bounds__(kTileX) void k_finalize1(FinArgs a) { finalize1_body(a, xcd_block(blockIdx.x, gridDim.x)); }
__global__ __launch_bounds__(kTileX) void k_finalize2(FinArgs a) {
  const int vb = xcd_block(blockIdx.x, gridDim.x);
  if (a.open_x | a.open_y) finalize2_body<false, true>(a, vb, DiagArgs{});
  else finalize2_body<false, false>(a, vb, DiagArgs{});
}

// rate arrays of resistive_B_term for the unit entry point.
__global__ __launch_bounds__(256) void k_rate(FinArgs a, int zero) {
  const int nx = a.nx, ny = a.ny;
  const int n = nx * ny;
  for (int o = blockIdx.x * blockDim.x + threadIdx.x; o < n; o += gridDim.x * blockDim.x) {
    const int j = o / nx, i = o - j * nx;
    const int jp = bc_wrap(j + 1, ny, a.open_y);
    const int ip = bc_wrap(i + 1, nx, a.open_x);
    if (zero) {
      a.bx_out[o] = -(0.0 - 0.0) * a.inv_dy;
      a.by_out[o] = (0.0 - 0.0) * a.inv_dx;
    } else {
      a.bx_out[o] = -(a.x[jp * nx + i] - a.x[j * nx + i]) * a.inv_dy;
      a.by_out[o] = (a.x[j * nx + ip] - a.x[j * nx + i]) * a.inv_dx;
    }
  }
}

// ---------------------------------------------------------------------------
// Diagnostics and layout helpers.
// ---------------------------------------------------------------------------

// CFL_condition_checker's wave speed (diagnostics.hpp:58-80) per cell, with
// Grid::computePrimitives' primitives; per-block max (NaN skipped exactly as
// the reference's `>` comparison skips it) and count of non-finite speeds.
__device__ __forceinline__ void cfl_body(const DiagArgs& a, int vb) {
  __shared__ double lds[16];
  const int nx = a.nx, ny = a.ny;
  const Tile t = tile_of(nx, ny, a.ty, a.tiles_y, vb, a.open_x);
  double m = 1e-8;
  double bad[1] = {0.0};
  if (t.valid) {
    auto at = [&](int col, int j) -> int { return wrap_row(j, ny, a.open_y) * nx + col; };
    double by_lo = a.s[5][at(t.ic, t.y0)];
    int o = t.y0 * nx + t.ic;
    double rho = a.s[0][o], mx = a.s[1][o], my = a.s[2][o], E = a.s[3][o];
    double bxl = a.s[4][o], bxr = a.s[4][t.y0 * nx + t.ip], by_hi = a.s[5][at(t.ic, t.y0 + 1)];
    for (int j = t.y0; j < t.y1; ++j) {
      const bool more = j + 1 < t.y1;
      // loads for the next row
      const int on = more ? (j + 1) * nx + t.ic : o;
      const double rho_n = a.s[0][on], mx_n = a.s[1][on], my_n = a.s[2][on], E_n = a.s[3][on];
      const double bxl_n = a.s[4][on];
      const double bxr_n = a.s[4][more ? (j + 1) * nx + t.ip : o];
      const double by_hi_n = more ? a.s[5][at(t.ic, j + 2)] : 0.0;
      const double bxc = (bxl + bxr) / 2.0;
      const double byc = (by_lo + by_hi) / 2.0;
      const double rr = (a.rho_raw && rho == kFloor) ? a.rho_raw[o] : rho;
      cfl_cell(rr, mx, my, E, bxc, byc, m, bad[0]);
      by_lo = by_hi; by_hi = by_hi_n;
      rho = rho_n; mx = mx_n; my = my_n; E = E_n; bxl = bxl_n; bxr = bxr_n;
      o = on;
    }
  }
  block_max(m, lds);
  block_sum<1>(bad, lds + 8);
  if (threadIdx.x == 0) {
    a.part[t.vb * 2 + 0] = m;
    a.part[t.vb * 2 + 1] = bad[0];
  }
}

__global__ __launch_bounds__(kTileX) void k_cfl(DiagArgs a) { cfl_body(a, xcd_block(blockIdx.x, gridDim.x)); }

// calculate_current_div_B (diagnostics.hpp:15-42): per-block sum of squares and max |div|.
__global__ __launch_bounds__(kTileX) void k_divb(DiagArgs a) {
  __shared__ double lds[16];
  const int nx = a.nx, ny = a.ny;
  const Tile t = tile_of(nx, ny, a.ty, a.tiles_y, xcd_block(blockIdx.x, gridDim.x), a.open_x);
  double ss[1] = {0.0};
  double m = 0.0;
  if (t.y0 < ny && t.valid) {
    double by_lo = a.s[5][t.y0 * nx + t.ic];
    for (int j = t.y0; j < t.y1; ++j) {
      const int jp = bc_wrap(j + 1, ny, a.open_y);
      const int o = j * nx + t.ic;
      const double by_hi = a.s[5][jp * nx + t.ic];
      const double xd = (a.s[4][j * nx + t.ip] - a.s[4][o]) / a.dx;
      const double yd = (by_hi - by_lo) / a.dy;
      by_lo = by_hi;
      const double tt = xd + yd;
      ss[0] += tt * tt;
      if (fabs(tt) > m) m = fabs(tt);
    }
  }
  block_sum<1>(ss, lds);
  block_max(m, lds + 8);
  if (threadIdx.x == 0) {
    a.part[t.vb * 2 + 0] = ss[0];
    a.part[t.vb * 2 + 1] = m;
  }
}

// Primitives of every cell, for mhd_download.
__global__ __launch_bounds__(256) void k_prims(DiagArgs a, double* vx, double* vy, double* p, double* bxc,
                                               double* byc) {
  const int nx = a.nx, ny = a.ny;
  const int n = nx * ny;
  for (int o = blockIdx.x * blockDim.x + threadIdx.x; o < n; o += gridDim.x * blockDim.x) {
    const int j = o / nx, i = o - j * nx;
    const int ip = bc_wrap(i + 1, nx, a.open_x);
    const int jp = bc_wrap(j + 1, ny, a.open_y);
    const double bx = (a.s[4][o] + a.s[4][j * nx + ip]) / 2.0;
    const double by = (a.s[5][o] + a.s[5][jp * nx + i]) / 2.0;
    double r, u, v, pp;
    primitives(a.s[0][o], a.s[1][o], a.s[2][o], a.s[3][o], bx, by, r, u, v, pp);
    vx[o] = u; vy[o] = v; p[o] = pp; bxc[o] = bx; byc[o] = by;
  }
}

// dst[q] = src[q] for q < nf (nf <= 6), n doubles each.
__global__ __launch_bounds__(256) void k_copy_fields(CopyArgs a, long long n) {
  for (long long o = blockIdx.x * (long long)blockDim.x + threadIdx.x; o < n; o += (long long)gridDim.x * blockDim.x) {
#pragma unroll
    for (int q = 0; q < 6; ++q)
      if (q < a.nf) a.dst[q][o] = a.src[q][o];
  }
}

// Fresh CG state (done = 0, counters zero) with the iteration limit.
__global__ void k_cg_reset(CgState* st, long long max_iters) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    CgState s{};
    s.max_iters = max_iters;
    *st = s;
  }
}

// Density floor applied to a freshly uploaded state (Grid.hpp:115).
__global__ __launch_bounds__(256) void k_floor_rho(double* rho, int n) {
  for (int o = blockIdx.x * blockDim.x + threadIdx.x; o < n; o += gridDim.x * blockDim.x)
    if (rho[o] < kFloor) rho[o] = kFloor;
}

__global__ __launch_bounds__(256) void k_floor_rho_raw(double* rho, double* raw, int n) {
  for (int o = blockIdx.x * blockDim.x + threadIdx.x; o < n; o += gridDim.x * blockDim.x) {
    const double r = rho[o];
    if (r <= kFloor) {
      raw[o] = r;
      if (r < kFloor) rho[o] = kFloor;
    }
  }
}

__global__ __launch_bounds__(256) void k_raw_reset(const double* rho, double* raw, int n) {
  for (int o = blockIdx.x * blockDim.x + threadIdx.x; o < n; o += gridDim.x * blockDim.x)
    if (rho[o] <= kFloor) raw[o] = rho[o];
}

__global__ __launch_bounds__(256) void k_copy(const double* __restrict__ src, double* __restrict__ dst, long long n) {
  for (long long o = blockIdx.x * (long long)blockDim.x + threadIdx.x; o < n; o += (long long)gridDim.x * blockDim.x)
    dst[o] = src[o];
}

// Streaming reference of a read/write mix: NI input and NO output arrays of
// n doubles, each element read once and written once (the outputs are the
// sum of the inputs plus the output's index), nontemporal loads and stores.
// A 1 GiB copy measured 4.2-6.2 TB/s across launch shapes (grids of 1 024
// to 65 536 blocks of 256, double or double2 per lane, cached or
// nontemporal), fastest on 65 536 blocks (5.9 TB/s cached, 6.2 TB/s
// nontemporal double2); this kernel takes that grid and nontemporal accesses
// with one double per lane, as the step's kernels access their fields.
template <int NI, int NO>
__global__ __launch_bounds__(256) void k_stream_mix(MixArgs a, long long n) {
  for (long long o = blockIdx.x * (long long)blockDim.x + threadIdx.x; o < n; o += (long long)gridDim.x * blockDim.x) {
    double v = 0.0;
#pragma unroll
    for (int q = 0; q < NI; ++q) v += __builtin_nontemporal_load(&a.in[q][o]);
#pragma unroll
    for (int q = 0; q < NO; ++q) __builtin_nontemporal_store(v + q, &a.out[q][o]);
  }
}
template __global__ void k_stream_mix<1, 1>(MixArgs, long long);
template __global__ void k_stream_mix<2, 1>(MixArgs, long long);
template __global__ void k_stream_mix<2, 2>(MixArgs, long long);
template __global__ void k_stream_mix<3, 2>(MixArgs, long long);
template __global__ void k_stream_mix<7, 2>(MixArgs, long long);
template __global__ void k_stream_mix<6, 12>(MixArgs, long long);
template __global__ void k_stream_mix<12, 6>(MixArgs, long long);


// ---------------------------------------------------------------------------
// Batched ensembles.  Each kernel below is the single-run kernel of the same
// name applied to every active run of a batch in one launch: the block reads
// its (run, block-within-run) job, forms the run's argument block from the
// pooled arrays and the run's geometry and driver state, and runs the
// single-run body.  A run whose driver has stopped returns at once.
// ---------------------------------------------------------------------------

template <int MODE>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void kb_stage(BatchStageArgs b) {
  if (b.ctl && b.ctl[kCtlHalt]) return;
  const int2 job = b.jobs[blockIdx.x];
  const RunState& rs = b.rs[job.x];
  if (!rs.active) return;
  const RunGeom& g = b.geom[job.x];
  StageArgs a;
  a.open_x = g.open & 1;
  a.open_y = (g.open >> 1) & 1;
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    a.s[q] = b.s[q] + g.base;
    a.u[q] = (MODE == 2) ? b.u[q] + g.base : nullptr;
    a.o1[q] = b.o1[q] + g.base;
    a.o2[q] = (MODE == 1) ? b.o2[q] + g.base : nullptr;
  }
  a.nx = g.nx;
  a.ny = g.ny;
  a.rows = g.rows;
  a.dx = g.dx;
  a.dy = g.dy;
  a.inv_dx = g.inv_dx;
  a.inv_dy = g.inv_dy;
  a.eta = g.eta;
  a.c_a = (MODE == 1) ? rs.coef[0] : 0.0;
  a.c_b = (MODE == 1) ? rs.coef[1] : rs.coef[2];
  a.c_g = (MODE == 2) ? rs.coef[3] : 0.0;
  a.rho_raw = (MODE == 2 && b.rho_raw) ? b.rho_raw + g.base : nullptr;
  stage_body<MODE>(a, job.y % g.strips, job.y / g.strips);
}
template __global__ void kb_stage<1>(BatchStageArgs);
template __global__ void kb_stage<2>(BatchStageArgs);

__device__ __forceinline__ CgArgs batch_cg_args(const BatchCgArgs& b, int run) {
  const RunGeom& g = b.geom[run];
  const RunState& rs = b.rs[run];
  CgArgs a;
  a.open_x = g.open & 1;
  a.open_y = (g.open >> 1) & 1;
  a.nx = g.nx;
  a.ny = g.ny;
  a.ty = g.ty;
  a.tiles_y = g.tiles_y;
  a.dx = g.dx;
  a.dy = g.dy;
  a.inv_dx = g.inv_dx;
  a.inv_dy = g.inv_dy;
  a.eta = g.eta;
  a.invd = rs.invd;
  a.diag = rs.diag;
  a.cx = rs.cx;
  a.cy = rs.cy;
  a.tol = b.tol;
  a.bx = b.bx + g.base;
  a.by = b.by + g.base;
  a.r = b.r + g.base;
  a.p0 = b.p0 + g.base;
  a.x = b.x + g.base;
  a.part_s = b.part_s + 3LL * g.part_base;
  a.part_a = b.part_a + g.part_base;
  a.part_b = b.part_b + 2LL * g.part_base;
  a.st = b.st + run;
  return a;
}

template <bool OPEN>
__global__ __launch_bounds__(kTileX) void kb_cg_start(BatchCgArgs b) {
  if (b.ctl && b.ctl[kCtlHalt]) return;
  const int2 job = b.jobs[xcd_block(blockIdx.x, gridDim.x)];
  if (b.st[job.x].done) return;            // stopped run, or eta = 0: no solve
  const CgArgs a = batch_cg_args(b, job.x);
  cg_start_body<true, OPEN>(a, nullptr, job.y);
}
template __global__ void kb_cg_start<false>(BatchCgArgs);
template __global__ void kb_cg_start<true>(BatchCgArgs);

// Start of a solve with the given right-hand side bin (pooled).
template <bool OPEN>
__global__ __launch_bounds__(kTileX) void kb_cg_start_rhs(BatchCgArgs b, const double* __restrict__ bin) {
  if (b.ctl && b.ctl[kCtlHalt]) return;
  const int2 job = b.jobs[xcd_block(blockIdx.x, gridDim.x)];
  if (b.st[job.x].done) return;
  const CgArgs a = batch_cg_args(b, job.x);
  cg_start_body<false, OPEN>(a, bin + b.geom[job.x].base, job.y);
}
template __global__ void kb_cg_start_rhs<false>(BatchCgArgs, const double*);
template __global__ void kb_cg_start_rhs<true>(BatchCgArgs, const double*);

template <bool OPEN>
__global__ __launch_bounds__(kTileX) void kb_cg_pass_a(BatchCgArgs b, const double* __restrict__ pold,
                                                       double* __restrict__ pnew, int it) {
  if (b.ctl && b.ctl[kCtlHalt]) return;
  const int2 job = b.jobs[xcd_block(blockIdx.x, gridDim.x)];
  if (b.st[job.x].done) return;
  const CgArgs a = batch_cg_args(b, job.x);
  const long long base = b.geom[job.x].base;
  cg_pass_a_body<OPEN>(a, pold + base, pnew + base, it, job.y, b.geom[job.x].tiles);
}
template __global__ void kb_cg_pass_a<false>(BatchCgArgs, const double*, double*, int);
template __global__ void kb_cg_pass_a<true>(BatchCgArgs, const double*, double*, int);

template <bool FIRST, bool OPEN>
__global__ __launch_bounds__(kTileX) void kb_cg_pass_b(BatchCgArgs b, const double* __restrict__ pv, int it) {
  if (b.ctl && b.ctl[kCtlHalt]) return;
  const int2 job = b.jobs[xcd_block(blockIdx.x, gridDim.x)];
  if (b.st[job.x].done) return;
  const CgArgs a = batch_cg_args(b, job.x);
  cg_pass_b_body<FIRST, OPEN>(a, pv + b.geom[job.x].base, it, job.y, b.geom[job.x].tiles);
}
template __global__ void kb_cg_pass_b<true, false>(BatchCgArgs, const double*, int);
template __global__ void kb_cg_pass_b<false, false>(BatchCgArgs, const double*, int);
template __global__ void kb_cg_pass_b<true, true>(BatchCgArgs, const double*, int);
template __global__ void kb_cg_pass_b<false, true>(BatchCgArgs, const double*, int);

__device__ __forceinline__ FinArgs batch_fin_args(const BatchFinArgs& b, int run) {
  const RunGeom& g = b.geom[run];
  FinArgs f;
  f.open_x = g.open & 1;
  f.open_y = (g.open >> 1) & 1;
  f.nx = g.nx;
  f.ny = g.ny;
  f.ty = g.ty;
  f.tiles_y = g.tiles_y;
  f.dx = g.dx;
  f.dy = g.dy;
  f.inv_dx = g.inv_dx;
  f.inv_dy = g.inv_dy;
  f.eta = g.eta;
  f.c = b.rs[run].coef[4];
  f.x = b.x + g.base;
  f.bx = b.bx + g.base;
  f.by = b.by + g.base;
  f.bx_out = b.bx_out + g.base;
  f.by_out = b.by_out + g.base;
  f.st = b.st + run;
  return f;
}

__global__ __launch_bounds__(kTileX) void kb_finalize1(BatchFinArgs b) {
  if (b.ctl && b.ctl[kCtlHalt]) return;
  const int2 job = b.jobs[xcd_block(blockIdx.x, gridDim.x)];
  if (!b.rs[job.x].active) return;
  finalize1_body(batch_fin_args(b, job.x), job.y);
}

// The host swaps the B arrays after this kernel for all runs, so a run that
// no longer steps copies its B across unchanged.
template <bool OPEN>
__global__ __launch_bounds__(kTileX) void kb_finalize2(BatchFinArgs b) {
  if (b.ctl && b.ctl[kCtlHalt]) return;
  const int2 job = b.jobs[xcd_block(blockIdx.x, gridDim.x)];
  const FinArgs f = batch_fin_args(b, job.x);
  if (b.rs[job.x].active) {
    finalize2_body<false, OPEN>(f, job.y, DiagArgs{});
    return;
  }
  const Tile t = tile_of(f.nx, f.ny, f.ty, f.tiles_y, job.y);
  if (!t.valid) return;
  for (int j = t.y0; j < t.y1; ++j) {
    const int o = j * f.nx + t.ic;
    f.bx_out[o] = f.bx[o];
    f.by_out[o] = f.by[o];
  }
}
template __global__ void kb_finalize2<false>(BatchFinArgs);
template __global__ void kb_finalize2<true>(BatchFinArgs);

// Per-tile CFL partials of every run that has not reached its end yet (the
// driver state still holds the previous step's activity; kb_dt decides).
__global__ __launch_bounds__(kTileX) void kb_cfl(BatchDiagArgs b) {
  if (b.ctl && b.ctl[kCtlHalt]) return;
  const int2 job = b.jobs[xcd_block(blockIdx.x, gridDim.x)];
  const RunGeom& g = b.geom[job.x];
  const RunState& rs = b.rs[job.x];
  if (!(rs.t < g.t_end) || (g.max_steps >= 0 && rs.steps >= g.max_steps)) return;
  DiagArgs d;
  d.open_x = g.open & 1;
  d.open_y = (g.open >> 1) & 1;
  d.nx = g.nx;
  d.ny = g.ny;
  d.ty = g.ty;
  d.tiles_y = g.tiles_y;
  d.dx = g.dx;
  d.dy = g.dy;
#pragma unroll
  for (int q = 0; q < 6; ++q) d.s[q] = b.s[q] + g.base;
  d.part = b.part + 2LL * g.part_base;
  d.rho_raw = b.rho_raw ? b.rho_raw + g.base : nullptr;
  cfl_body(d, job.y);
}

// DiagArgs of one run of a batch (its tiles, its slice of the pooled fields
// and partials).
__device__ __forceinline__ DiagArgs batch_diag_args(const BatchDiagArgs& b, int run) {
  const RunGeom& g = b.geom[run];
  DiagArgs d;
  d.open_x = g.open & 1;
  d.open_y = (g.open >> 1) & 1;
  d.nx = g.nx;
  d.ny = g.ny;
  d.ty = g.ty;
  d.tiles_y = g.tiles_y;
  d.dx = g.dx;
  d.dy = g.dy;
#pragma unroll
  for (int q = 0; q < 6; ++q) d.s[q] = b.s[q] + g.base;
  d.part = b.part + 2LL * g.part_base;
  d.rho_raw = b.rho_raw ? b.rho_raw + g.base : nullptr;
  return d;
}

// kb_finalize2 that also forms the next step's CFL partials of every active
// run from the new state (rho, mx, my, E of U3*, the new B in registers), so
// the next step needs no kb_cfl.  Runs that did not step copy their B and
// keep no partials (a run that did not step this time does not step next).
template <bool OPEN>
__global__ __launch_bounds__(kTileX) void kb_finalize2_cfl(BatchFinArgs b, BatchDiagArgs dg) {
  if (b.ctl && b.ctl[kCtlHalt]) return;
  const int2 job = b.jobs[xcd_block(blockIdx.x, gridDim.x)];
  const FinArgs f = batch_fin_args(b, job.x);
  if (b.rs[job.x].active) {
    finalize2_body<true, OPEN>(f, job.y, batch_diag_args(dg, job.x));
    return;
  }
  const Tile t = tile_of(f.nx, f.ny, f.ty, f.tiles_y, job.y);
  if (!t.valid) return;
  for (int j = t.y0; j < t.y1; ++j) {
    const int o = j * f.nx + t.ic;
    f.bx_out[o] = f.bx[o];
    f.by_out[o] = f.by[o];
  }
}
template __global__ void kb_finalize2_cfl<false>(BatchFinArgs, BatchDiagArgs);
template __global__ void kb_finalize2_cfl<true>(BatchFinArgs, BatchDiagArgs);

// Driver step of every run (src/integrator_main.cpp:76-100 of the reference,
// as mhd_run does it for one run): activity, CFL time step with the Courant
// number of the run (CFL_condition_checker, diagnostics.hpp:52-86), the clamp
// of the last step to t_end, the operator rebuild when dt changed by more than
// 1e-10 relative (alpha = dt eta gamma), then t and the step count advance.
// Also resets both CG states of the run.  One block per run.
__global__ __launch_bounds__(256) void kb_dt(const RunGeom* __restrict__ geom, RunState* __restrict__ rs,
                                             const double* __restrict__ part, CgState* __restrict__ st0,
                                             CgState* __restrict__ st1, int nruns, StepFactors f,
                                             int halt_on_nonfinite, int* __restrict__ ctl) {
  const double a_im4 = f.im4;
  __shared__ double lds[16];
  const int run = blockIdx.x;
  if (run >= nruns) return;
  if (ctl && ctl[kCtlHalt]) return;          // halted group: nothing moves until the host resumes
  if (ctl && run == 0 && threadIdx.x == 0) ctl[kCtlSteps] += 1;
  const RunGeom& g = geom[run];
  RunState s = rs[run];
  if (s.halt) return;                        // halted resident loop: nothing moves until the host resumes
  const bool active = (s.t < g.t_end) && (g.max_steps < 0 || s.steps < g.max_steps);
  double m = 1e-8;
  double bad[1] = {0.0};
  if (active) {
    for (int b = threadIdx.x; b < g.tiles; b += blockDim.x) {
      m = max_ignoring_nan(m, part[2LL * (g.part_base + b)]);
      bad[0] += part[2LL * (g.part_base + b) + 1];
    }
  }
  block_max(m, lds);
  block_sum<1>(bad, lds + 8);
  if (threadIdx.x != 0) return;
  CgState c{};
  c.max_iters = g.max_iters;
  if (!active) {
    s.active = 0;
    c.done = 3;
  } else {
    if (halt_on_nonfinite && g.dt_fixed <= 0.0 && bad[0] > 0.0) {
      s.nonfinite = 1;                       // the step is not taken (mhd_cfl_dt fails before it)
      if (ctl) ctl[kCtlHalt] = kHaltNonfinite;
      else s.halt = kHaltNonfinite;
      rs[run] = s;
      return;
    }
    const double h = (g.dy < g.dx) ? g.dy : g.dx;
    const double dt_res = (g.eta != 0.0) ? h * h / (4.0 * g.eta) : INFINITY;
    const double dt_hyp = (g.cfl / 0.9) * h / m;
    const double dt_cfl = ((dt_res < dt_hyp) ? dt_res : dt_hyp) * 0.9;
    const double dt0 = (g.dt_fixed > 0.0) ? g.dt_fixed : dt_cfl;
    if (s.steps == 0 && !(s.prev_dt > 0.0)) {   // operator of the first dt, before the loop (unless preset)
      s.alpha = (dt0 * g.eta) * a_im4;
      s.prev_dt = dt0;
    }
    double dt = dt0;
    if (s.t + dt > g.t_end) dt = g.t_end - s.t;
    if (g.eta > 0.0 && fabs(dt - s.prev_dt) > 1e-10 * dt) s.alpha = (dt * g.eta) * a_im4;
    s.prev_dt = dt;
    s.dt = dt;
    s.coef[0] = dt * f.ex3;
    s.coef[1] = dt * f.ex6;
    s.coef[2] = dt * f.ex7;
    s.coef[3] = dt * f.im7;
    s.coef[4] = dt * f.im4;
    s.cx = s.alpha / (g.dx * g.dx);
    s.cy = s.alpha / (g.dy * g.dy);
    s.diag = 1 + 2 * (s.cx + s.cy);
    s.invd = 1.0 / s.diag;
    s.active = 1;
    if (g.dt_fixed <= 0.0 && bad[0] > 0.0) s.nonfinite = 1;
    s.t += dt;
    s.steps += 1;
    if (g.eta == 0.0) c.done = 3;            // no implicit solve
  }
  rs[run] = s;
  st0[run] = c;
  st1[run] = c;
}

// ---------------------------------------------------------------------------
// Single-run device-resident driver loop (mhd_run_resident).  kb_dt (one run)
// writes the step's dt, the operator coefficients and the run's activity into
// a RunState; the kernels below are the plain kernels with those values read
// from it, so the host enqueues many steps without reading anything back.
// Each returns at once when the run has stopped (t >= t_end or the step
// limit) or the loop has halted.  The host launches a fixed number of CG
// iterations per solve (the passes of a finished solve return at once); a
// finaliser that finds its solve unfinished halts the loop, and the host
// finishes that step itself.
// ---------------------------------------------------------------------------

__device__ __forceinline__ bool res_stopped(const RunState* rs) { return !rs->active || rs->halt != 0; }

template <int MODE>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(MHD_STAGE_WAVES))) void k_stage_res(
    StageArgs a, const RunState* __restrict__ rs) {
  if (res_stopped(rs)) return;
  StageArgs b = a;
  b.c_a = (MODE == 1) ? rs->coef[0] : 0.0;
  b.c_b = (MODE == 1) ? rs->coef[1] : rs->coef[2];
  b.c_g = (MODE == 2) ? rs->coef[3] : 0.0;
  const int2 t = stage_tile(a);
  stage_body<MODE>(b, t.x, t.y);
}
template __global__ void k_stage_res<1>(StageArgs, const RunState*);
template __global__ void k_stage_res<2>(StageArgs, const RunState*);

__device__ __forceinline__ CgArgs res_cg_args(const CgArgs& a, const RunState* rs) {
  CgArgs b = a;
  b.cx = rs->cx;
  b.cy = rs->cy;
  b.diag = rs->diag;
  b.invd = rs->invd;
  return b;
}

template <bool OPEN>
__global__ __launch_bounds__(kTileX) void k_cg_start_res(CgArgs a, const RunState* __restrict__ rs) {
  if (res_stopped(rs)) return;
  cg_start_body<true, OPEN>(res_cg_args(a, rs), nullptr, xcd_block(blockIdx.x, gridDim.x));
}
template __global__ void k_cg_start_res<false>(CgArgs, const RunState*);
template __global__ void k_cg_start_res<true>(CgArgs, const RunState*);

template <bool OPEN>
__global__ __launch_bounds__(kTileX) void k_cg_pass_a_res(CgArgs a, const RunState* __restrict__ rs,
                                                          const double* __restrict__ pold, double* __restrict__ pnew,
                                                          int it) {
  if (res_stopped(rs)) return;
  cg_pass_a_body<OPEN>(res_cg_args(a, rs), pold, pnew, it, xcd_block(blockIdx.x, gridDim.x), gridDim.x);
}
template __global__ void k_cg_pass_a_res<false>(CgArgs, const RunState*, const double*, double*, int);
template __global__ void k_cg_pass_a_res<true>(CgArgs, const RunState*, const double*, double*, int);

template <bool FIRST, bool OPEN>
__global__ __launch_bounds__(kTileX) void k_cg_pass_b_res(CgArgs a, const RunState* __restrict__ rs,
                                                          const double* __restrict__ pv, int it) {
  if (res_stopped(rs)) return;
  cg_pass_b_body<FIRST, OPEN>(res_cg_args(a, rs), pv, it, xcd_block(blockIdx.x, gridDim.x), gridDim.x);
}
template __global__ void k_cg_pass_b_res<true, false>(CgArgs, const RunState*, const double*, int);
template __global__ void k_cg_pass_b_res<false, false>(CgArgs, const RunState*, const double*, int);
template __global__ void k_cg_pass_b_res<true, true>(CgArgs, const RunState*, const double*, int);
template __global__ void k_cg_pass_b_res<false, true>(CgArgs, const RunState*, const double*, int);

// Log entry of the step just finished (block 0, thread 0 of the caller).
__device__ __forceinline__ void write_step_log(const RunState* rs, const CgState* st, StepLog* log, int log_cap) {
  StepLog e;
  e.dt = rs->dt;
  e.t = rs->t;
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const CgState& s = st[k];
    e.iters[k] = s.iters;
    e.spmv[k] = s.spmv;
    e.err[k] = (s.done == 1 || s.done == 2) ? sqrt(s.rn2 / s.rhs_norm2) : 0.0;
  }
  log[(rs->steps - 1) % log_cap] = e;
}

// Every block of a finaliser takes the same decision from its solve's state:
// unfinished (or at the iteration limit with fail_is_error) halts the loop
// before any field changes.
__device__ __forceinline__ bool halt_unfinished(const CgState* st, RunState* rs, int slot, int fail_is_error) {
  const int done = st->done;
  if (done == 0 || (done == 2 && fail_is_error)) {
    if (blockIdx.x == 0 && threadIdx.x == 0)
      rs->halt = (done == 0) ? (slot == 0 ? kHaltSolve1 : kHaltSolve2) : (slot == 0 ? kHaltCgLimit1 : kHaltCgLimit2);
    return true;
  }
  return false;
}

// Finaliser of solve 1: B2 = B2* + c rate(x).
__global__ __launch_bounds__(kTileX) void k_finalize1_res(FinArgs a, RunState* __restrict__ rs, int fail_is_error) {
  if (res_stopped(rs)) return;
  if (halt_unfinished(a.st, rs, 0, fail_is_error)) return;
  FinArgs f = a;
  f.c = rs->coef[4];
  finalize1_body(f, xcd_block(blockIdx.x, gridDim.x));
}

// Finaliser of solve 2 (the final recombination), the end of a resistive
// step: the step's log entry and, with CFL, the CFL partials of the new
// state (d: its rho, mx, my, E and the partials buffer) for the next kb_dt.
template <bool CFL>
__global__ __launch_bounds__(kTileX) void k_finalize2_res(FinArgs a, DiagArgs d, RunState* __restrict__ rs,
                                                          const CgState* __restrict__ st, StepLog* __restrict__ log,
                                                          int log_cap, int fail_is_error) {
  if (res_stopped(rs)) return;
  if (halt_unfinished(a.st, rs, 1, fail_is_error)) return;
  if (blockIdx.x == 0 && threadIdx.x == 0) write_step_log(rs, st, log, log_cap);
  FinArgs f = a;
  f.c = rs->coef[4];
  const int vb = xcd_block(blockIdx.x, gridDim.x);
  if (a.open_x | a.open_y) finalize2_body<CFL, true>(f, vb, d);
  else finalize2_body<CFL, false>(f, vb, d);
}
template __global__ void k_finalize2_res<false>(FinArgs, DiagArgs, RunState*, const CgState*, StepLog*, int, int);
template __global__ void k_finalize2_res<true>(FinArgs, DiagArgs, RunState*, const CgState*, StepLog*, int, int);

// End of a step without a resistive solve: the step's log entry (block 0)
// and, with `partials`, the CFL partials of the new state for the next kb_dt.
__global__ __launch_bounds__(kTileX) void k_cfl_res(DiagArgs a, const RunState* __restrict__ rs,
                                                    const CgState* __restrict__ st, StepLog* __restrict__ log,
                                                    int log_cap, int partials) {
  if (res_stopped(rs)) return;
  if (blockIdx.x == 0 && threadIdx.x == 0) write_step_log(rs, st, log, log_cap);
  if (partials) cfl_body(a, xcd_block(blockIdx.x, gridDim.x));
}

__global__ void k_clear_halt(RunState* rs) {
  if (blockIdx.x == 0 && threadIdx.x == 0) rs->halt = 0;
}

__global__ __launch_bounds__(256) void kb_check_solve(const RunState* __restrict__ rs, const CgState* __restrict__ st,
                                                      int nruns, int slot, int fail_is_error, int* __restrict__ ctl,
                                                      int* __restrict__ iters) {
  if (ctl[kCtlHalt]) return;
  __shared__ int code;
  if (threadIdx.x == 0) code = 0;
  __syncthreads();
  for (int r = threadIdx.x; r < nruns; r += blockDim.x) {
    if (!rs[r].active) continue;
    const int done = st[r].done;
    if (done == 0) atomicMax(&code, 2);
    else if (done == 2 && fail_is_error) atomicMax(&code, 1);
    iters[2 * r + slot] = st[r].iters;
  }
  __syncthreads();
  if (threadIdx.x == 0 && code != 0)
    ctl[kCtlHalt] = (code == 2) ? (slot == 0 ? kHaltSolve1 : kHaltSolve2) : (slot == 0 ? kHaltCgLimit1 : kHaltCgLimit2);
}

__global__ void kb_ctl_reset(int* ctl) {
  if (blockIdx.x == 0 && threadIdx.x < 4) ctl[threadIdx.x] = 0;
}

}  // namespace mhd
