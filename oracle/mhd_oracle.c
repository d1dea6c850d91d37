/*
 * mhd_oracle.c -- CPU restatement of the reference MHD step (TEST INFRASTRUCTURE).
 *
 * Parity checker for the HIP path; not part of the product.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 *
 * Every function below follows the reference file:line named in its comment
 * (paths relative to the reference repository).  Expressions keep the
 * reference's operation order so that a strict build (-ffp-contract=off, no
 * fast-math) reproduces the reference headers bit for bit where they compile.
 *
 * How each part is pinned (tests/ holds the checks):
 *   grid, ghost fill, primitives, minmod, MUSCL, fast speed, HLL x/y, CFL, div(B):
 *       bitwise against the reference's own Grid.hpp / MUSCL.hpp / HLL.hpp /
 *       diagnostics.hpp, compiled from their sources into oracle/_ref
 *       (tests/test_oracle_ref.py), plus closed-form flux cases.
 *   flux assembly, upwind CT, time derivatives (solvers.hpp, which needs
 *       Eigen and cannot be compiled here): uniform state gives zero RHS,
 *       discrete div(dB/dt) == 0, conservation of mass/momentum/energy,
 *       mirror/rotation symmetry, y-independent data reduces to 1D MUSCL-HLL.
 *   resistive operator + CG: dense solve of (I - alpha L) with numpy on small
 *       grids, Fourier-mode closed form 1/(1 - alpha*lambda).
 *   IMEX combination: per-mode pure-diffusion factor
 *       R(z) = (1 + (1-2g) z) / (1 - g z)^2, second-order temporal convergence.
 *   Eigen's exact dot-product rounding order: parity unpinned (Eigen absent);
 *       the dense-solve and closed-form checks pin the mathematics.
 */
#include "mhd_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define GAMMA_GAS (5.0 / 3.0)   /* Grid.hpp:104, HLL.hpp:31,62,85,108, diagnostics.hpp:48 */
#define PI_REF 3.14159265359    /* integrator_main.cpp:36, benchmarks.hpp:88,190,245,306,363 */

struct og_grid {
  int nx, ny, ng, nxt, nyt;
  double dx, dy, eta;
  int bc_x, bc_y;   /* 0 periodic (the reference), 1 outflow (zero-gradient ghost rule) */
  double* f[OG_NFIELDS];
};

#define RHO f[OG_RHO]
#define MXF f[OG_MX]
#define MYF f[OG_MY]
#define EF f[OG_E]
#define BXF f[OG_BXF]
#define BYF f[OG_BYF]
#define VXF f[OG_VX]
#define VYF f[OG_VY]
#define PF f[OG_P]
#define BXC f[OG_BXC]
#define BYC f[OG_BYC]
#define EZF f[OG_EZ]

/* ReconstructedValues order, MUSCL.hpp:53-58 */
enum {
  RV_RHO_XL, RV_RHO_XR, RV_RHO_YL, RV_RHO_YR,
  RV_VX_XL, RV_VX_XR, RV_VX_YL, RV_VX_YR,
  RV_VY_XL, RV_VY_XR, RV_VY_YL, RV_VY_YR,
  RV_P_XL, RV_P_XR, RV_P_YL, RV_P_YR,
  RV_BX_YL, RV_BX_YR, RV_BY_XL, RV_BY_XR, RV_COUNT
};
/* flux arrays, solvers.hpp:120-141 */
enum { FL_RHO_H, FL_RHO_V, FL_MX_H, FL_MX_V, FL_MY_H, FL_MY_V, FL_E_H, FL_E_V,
       FL_BX_H, FL_BX_V, FL_BY_H, FL_BY_V, FL_COUNT };

struct og_ws {
  int nx, ny, nxt, nyt, ncell, nphys;
  double* rv[RV_COUNT];
  double* fl[FL_COUNT];
  double* f1[6]; double* f2[6];           /* rho, mx, my, E, bxf, byf */
  double* g1[2]; double* g2[2]; double* g3[2];
  double* solved1[2]; double* solved2[2];
  double* corner_e;                        /* electric_field_at_corners */
  double* ez_compact;                      /* Ez_corner_no_padding (N) */
  double* potential;                       /* magnetic_potential ((nxt+1)(nyt+1)) */
  /* CSR of A, rows = columns of the reference's column-major matrix */
  int* a_ptr; int* a_col; double* a_val; double a_invdiag_scratch;
  double* invdiag;
  double cg_tol; long long cg_maxit;
  /* CG scratch */
  double *r, *p, *z, *tmp, *xs;
  og_grid* inter;   /* grid_intermediary (integrator_main.cpp:41) */
  og_grid* whole;   /* grid_whole_step   (integrator_main.cpp:42) */
};

static double* dalloc(size_t n) {
  double* a = (double*)calloc(n ? n : 1, sizeof(double));
  return a;
}

/* ---------------- Grid (include/Grid.hpp) ---------------- */

/* Grid::Grid, Grid.hpp:29-54: 12 padded arrays, nxt = nx + 2 ng */
og_grid* og_grid_create(int nx, int ny, double dx, double dy, double eta) {
  og_grid* g = (og_grid*)calloc(1, sizeof(og_grid));
  g->nx = nx; g->ny = ny; g->ng = 2;
  g->nxt = nx + 4; g->nyt = ny + 4;
  g->dx = dx; g->dy = dy; g->eta = eta;
  size_t n = (size_t)g->nxt * g->nyt;
  for (int k = 0; k < OG_NFIELDS; ++k) g->f[k] = dalloc(n);
  return g;
}

void og_grid_destroy(og_grid* g) {
  if (!g) return;
  for (int k = 0; k < OG_NFIELDS; ++k) free(g->f[k]);
  free(g);
}

double* og_field(og_grid* g, int which) { return (which >= 0 && which < OG_NFIELDS) ? g->f[which] : NULL; }

void og_grid_set_eta(og_grid* g, double eta) { g->eta = eta; }

/* Boundary of each axis: 0 periodic (the reference's only one), 1 outflow.
 * The reference names the missing piece (Grid.hpp:141-142): a non-periodic
 * run needs another ghost-fill rule.  Outflow here is the zero-gradient rule:
 *   - conserved and face fields: ghosts take the value of the nearest cell
 *     (face) inside, wherever the reference fills ghosts;
 *   - cell-centred B of every cell, ghosts included, is the mean of its two
 *     ghost-filled faces (computeCellB, not a copy);
 *   - the CT update of the last face line uses E_z of the boundary corner
 *     line, which cell_corner_electric_fields computes over [-1, n] anyway;
 *   - corner quantities of resistive_B_term (eta J and the CG unknown) past
 *     the last corner take the last corner's value, and generate_A_matrix
 *     folds a clamped neighbour into the diagonal (a Neumann Laplacian).
 * This is the rule the library's kernels implement. */
void og_grid_set_bc(og_grid* g, int bc_x, int bc_y) { g->bc_x = bc_x; g->bc_y = bc_y; }

static inline int clamp_index(int k, int n) { return k < 0 ? 0 : (k >= n ? n - 1 : k); }

/* Grid copy assignment (used at solvers.hpp:493,501-502 and integrator_main.cpp:43-44) */
void og_grid_copy(og_grid* dst, const og_grid* src) {
  size_t n = (size_t)src->nxt * src->nyt;
  dst->nx = src->nx; dst->ny = src->ny; dst->ng = src->ng; dst->nxt = src->nxt; dst->nyt = src->nyt;
  dst->dx = src->dx; dst->dy = src->dy; dst->eta = src->eta;
  dst->bc_x = src->bc_x; dst->bc_y = src->bc_y;
  for (int k = 0; k < OG_NFIELDS; ++k) memcpy(dst->f[k], src->f[k], n * sizeof(double));
}

static og_grid* grid_clone(const og_grid* src) {
  og_grid* g = (og_grid*)calloc(1, sizeof(og_grid));
  size_t n = (size_t)src->nxt * src->nyt;
  *g = *src;
  for (int k = 0; k < OG_NFIELDS; ++k) {
    g->f[k] = (double*)malloc(n * sizeof(double));
    memcpy(g->f[k], src->f[k], n * sizeof(double));
  }
  return g;
}

static inline int IDX(const og_grid* g, int i, int j) { return j * g->nxt + i; }   /* Grid.hpp:58 */

/* Grid::fillOneField, Grid.hpp:143-161 */
static void fill_one_field(og_grid* g, double* f) {
  const int nx = g->nx, ny = g->ny, ng = g->ng, nxt = g->nxt, nyt = g->nyt;
  for (int j = ng; j < nyt - ng; ++j) {
    for (int gg = 0; gg < ng; ++gg) {
      f[IDX(g, gg, j)] = g->bc_x ? f[IDX(g, ng, j)] : f[IDX(g, nx + gg, j)];
      f[IDX(g, nx + ng + gg, j)] = g->bc_x ? f[IDX(g, nx + ng - 1, j)] : f[IDX(g, ng + gg, j)];
    }
  }
  for (int i = 0; i < nxt; ++i) {
    for (int gg = 0; gg < ng; ++gg) {
      f[IDX(g, i, gg)] = g->bc_y ? f[IDX(g, i, ng)] : f[IDX(g, i, ny + gg)];
      f[IDX(g, i, ny + ng + gg)] = g->bc_y ? f[IDX(g, i, ny + ng - 1)] : f[IDX(g, i, ng + gg)];
    }
  }
}

/* Ghosts of a cell-centred B array after og_compute_cell_b on an outflow
 * axis: computeCellB already formed every ghost cell from its faces except
 * the last padded column (row), which takes its neighbour's value; a
 * periodic axis copies as fill_one_field does. */
static void fill_cell_b(og_grid* g, double* f) {
  const int nx = g->nx, ny = g->ny, ng = g->ng, nxt = g->nxt, nyt = g->nyt;
  for (int j = ng; j < nyt - ng; ++j) {
    for (int gg = 0; gg < ng; ++gg) {
      if (g->bc_x) {
        f[IDX(g, nxt - 1, j)] = f[IDX(g, nxt - 2, j)];
      } else {
        f[IDX(g, gg, j)] = f[IDX(g, nx + gg, j)];
        f[IDX(g, nx + ng + gg, j)] = f[IDX(g, ng + gg, j)];
      }
    }
  }
  for (int i = 0; i < nxt; ++i) {
    for (int gg = 0; gg < ng; ++gg) {
      if (g->bc_y) {
        f[IDX(g, i, nyt - 1)] = f[IDX(g, i, nyt - 2)];
      } else {
        f[IDX(g, i, gg)] = f[IDX(g, i, ny + gg)];
        f[IDX(g, i, ny + ng + gg)] = f[IDX(g, i, ng + gg)];
      }
    }
  }
  if (g->bc_x) for (int j = 0; j < nyt; ++j) f[IDX(g, nxt - 1, j)] = f[IDX(g, nxt - 2, j)];
}

/* Grid::computeCellB, Grid.hpp:82-97 */
void og_compute_cell_b(og_grid* g) {
  for (int j = 0; j < g->nyt - 1; ++j) {
    for (int i = 0; i < g->nxt - 1; ++i) {
      g->BXC[IDX(g, i, j)] = (g->BXF[IDX(g, i, j)] + g->BXF[IDX(g, i + 1, j)]) / 2.0;
      g->BYC[IDX(g, i, j)] = (g->BYF[IDX(g, i, j)] + g->BYF[IDX(g, i, j + 1)]) / 2.0;
    }
  }
}

/* Grid::fillGhostPeriodic, Grid.hpp:64-80, with the grid's boundary rule
 * (identical to the reference's on periodic axes) */
void og_fill_ghost(og_grid* g) {
  fill_one_field(g, g->RHO);
  fill_one_field(g, g->MXF);
  fill_one_field(g, g->MYF);
  fill_one_field(g, g->EF);
  fill_one_field(g, g->BXF);
  fill_one_field(g, g->BYF);
  og_compute_cell_b(g);
  if (g->bc_x || g->bc_y) {
    fill_cell_b(g, g->BXC);
    fill_cell_b(g, g->BYC);
  } else {
    fill_one_field(g, g->BXC);
    fill_one_field(g, g->BYC);
  }
}

/* the reference's name; periodic grids are filled exactly as Grid.hpp:64-80 */
void og_fill_ghost_periodic(og_grid* g) { og_fill_ghost(g); }

/* Grid::computePrimitives, Grid.hpp:100-136.  Velocities use the density
 * before the floor; the floor then mutates rho; the E re-sync block at
 * :127-131 computes values it never stores, so it has no effect. */
void og_compute_primitives(og_grid* g) {
  const double gamma = GAMMA_GAS;
  for (int j = 0; j < g->nyt; ++j) {
    for (int i = 0; i < g->nxt; ++i) {
      int id = IDX(g, i, j);
      g->VXF[id] = g->MXF[id] / g->RHO[id];
      g->VYF[id] = g->MYF[id] / g->RHO[id];
      if (g->RHO[id] < 1e-8) g->RHO[id] = 1e-8;
      double kinE = 0.5 * g->RHO[id] * (g->VXF[id] * g->VXF[id] + g->VYF[id] * g->VYF[id]);
      double magE = (g->BXC[id] * g->BXC[id] + g->BYC[id] * g->BYC[id]) / 2;
      double intE = g->EF[id] - kinE - magE;
      g->PF[id] = (gamma - 1) * intE;
      if (g->PF[id] < 1e-8) g->PF[id] = 1e-8;
    }
  }
}

/* ---------------- MUSCL (include/MUSCL.hpp) ---------------- */

/* minmod, MUSCL.hpp:11-17 (std::abs on double) */
double og_minmod(double a, double b) {
  if (a * b > 0.0) return (fabs(a) < fabs(b)) ? a : b;
  return 0.0;
}

/* MUSCLReconstructOne, MUSCL.hpp:26-47.  dir 0 = x, 1 = y */
void og_muscl_one(const double* F, const og_grid* g, int i, int j, int dir, double* L, double* R) {
  double leftSlope, rightSlope;
  if (dir == 0) {
    leftSlope = F[IDX(g, i, j)] - F[IDX(g, i - 1, j)];
    rightSlope = F[IDX(g, i + 1, j)] - F[IDX(g, i, j)];
  } else {
    leftSlope = F[IDX(g, i, j)] - F[IDX(g, i, j - 1)];
    rightSlope = F[IDX(g, i, j + 1)] - F[IDX(g, i, j)];
  }
  double slope = og_minmod(leftSlope, rightSlope);
  *L = F[IDX(g, i, j)] - 0.5 * slope;
  *R = F[IDX(g, i, j)] + 0.5 * slope;
}

/* MUSCLReconstructAll, MUSCL.hpp:82-119: cells [-1, n] in physical coordinates */
void og_muscl_all(const og_grid* g, og_ws* w) {
  const double* fields[4] = {g->RHO, g->VXF, g->VYF, g->PF};
  for (int j = g->ng - 1; j <= g->ny + g->ng; j++) {
    for (int i = g->ng - 1; i <= g->nx + g->ng; i++) {
      int id = IDX(g, i, j);
      for (int a = 0; a < 4; a++) {
        double L, R;
        og_muscl_one(fields[a], g, i, j, 0, &L, &R);
        w->rv[4 * a + 0][id] = L; w->rv[4 * a + 1][id] = R;
        og_muscl_one(fields[a], g, i, j, 1, &L, &R);
        w->rv[4 * a + 2][id] = L; w->rv[4 * a + 3][id] = R;
      }
      double L, R;
      og_muscl_one(g->BXC, g, i, j, 1, &L, &R);
      w->rv[RV_BX_YL][id] = L; w->rv[RV_BX_YR][id] = R;
      og_muscl_one(g->BYC, g, i, j, 0, &L, &R);
      w->rv[RV_BY_XL][id] = L; w->rv[RV_BY_XR][id] = R;
    }
  }
}

/* ---------------- HLL (include/HLL.hpp) ---------------- */
/* PrimitiveState / ConservedState / FluxVector all use the order
 * (rho, vx|mx, vy|my, p|E, bx, by), HLL.hpp:12-58. */

static void from_primitive(const double w[6], double u[6]) {   /* HLL.hpp:31-47 */
  const double gamma = GAMMA_GAS;
  u[0] = w[0];
  u[1] = w[0] * w[1];
  u[2] = w[0] * w[2];
  double v_sq = w[1] * w[1] + w[2] * w[2];
  double b_sq = w[4] * w[4] + w[5] * w[5];
  u[3] = w[3] / (gamma - 1.0) + 0.5 * w[0] * v_sq + 0.5 * b_sq;
  u[4] = w[4];
  u[5] = w[5];
}

static void convert_primitive_x(const double p[6], double F[6]) {   /* HLL.hpp:61-81 */
  const double gamma = 5.0 / 3.0;
  double v_sq = p[1] * p[1] + p[2] * p[2];
  double B_sq = p[4] * p[4] + p[5] * p[5];
  double P_tot = p[3] + 0.5 * B_sq;
  double v_dot_B = p[1] * p[4] + p[2] * p[5];
  double E = p[3] / (gamma - 1.0) + 0.5 * (p[0] * v_sq + B_sq);
  F[0] = p[0] * p[1];
  F[1] = p[0] * p[1] * p[1] + P_tot - p[4] * p[4];
  F[2] = p[0] * p[1] * p[2] - p[4] * p[5];
  F[3] = (E + P_tot) * p[1] - p[4] * v_dot_B;
  F[4] = 0.0;
  F[5] = p[1] * p[5] - p[2] * p[4];
}

static void convert_primitive_y(const double p[6], double F[6]) {   /* HLL.hpp:84-104 */
  const double gamma = 5.0 / 3.0;
  double v_sq = p[1] * p[1] + p[2] * p[2];
  double B_sq = p[4] * p[4] + p[5] * p[5];
  double P_tot = p[3] + 0.5 * B_sq;
  double v_dot_B = p[1] * p[4] + p[2] * p[5];
  double E = p[3] / (gamma - 1.0) + 0.5 * (p[0] * v_sq + B_sq);
  F[0] = p[0] * p[2];
  F[2] = p[0] * p[2] * p[2] + P_tot - p[5] * p[5];
  F[1] = p[0] * p[1] * p[2] - p[4] * p[5];
  F[3] = (E + P_tot) * p[2] - p[5] * v_dot_B;
  F[5] = 0.0;
  F[4] = p[2] * p[4] - p[1] * p[5];
}

static inline double dmin(double a, double b) { return (b < a) ? b : a; }   /* std::min */
static inline double dmax(double a, double b) { return (a < b) ? b : a; }   /* std::max */

/* getFastWaveSpeed, HLL.hpp:107-122 */
double og_fast_speed(const double p[6], double bn) {
  const double gamma = 5.0 / 3.0;
  double B_sq = p[4] * p[4] + p[5] * p[5];
  double sound_v_sq = gamma * p[3] / p[0];
  double tot_alfven_v_sq = B_sq / p[0];
  double norm_alfven_v_sq = bn * bn / p[0];
  double sum = tot_alfven_v_sq + sound_v_sq;
  double term = sum * sum - 4.0 * sound_v_sq * norm_alfven_v_sq;
  return sqrt(0.5 * (sum + sqrt(dmax(0.0, term))));
}

/* computeHLLFluxX, HLL.hpp:126-172 */
void og_hll_x(const double L[6], const double R[6], double out[6]) {
  double FL[6], FR[6];
  convert_primitive_x(L, FL);
  convert_primitive_x(R, FR);
  double cL = og_fast_speed(L, L[4]);
  double cR = og_fast_speed(R, R[4]);
  double sL = dmin(L[1] - cL, R[1] - cR);
  double sR = dmax(L[1] + cL, R[1] + cR);
  if (sL >= 0) { memcpy(out, FL, sizeof FL); return; }
  if (sR <= 0) { memcpy(out, FR, sizeof FR); return; }
  double UL[6], UR[6];
  from_primitive(L, UL);
  from_primitive(R, UR);
  double inv_S = 1.0 / (sR - sL);
  out[0] = (sR * FL[0] - sL * FR[0] + sL * sR * (UR[0] - UL[0])) * inv_S;
  out[1] = (sR * FL[1] - sL * FR[1] + sL * sR * (UR[1] - UL[1])) * inv_S;
  out[2] = (sR * FL[2] - sL * FR[2] + sL * sR * (UR[2] - UL[2])) * inv_S;
  out[3] = (sR * FL[3] - sL * FR[3] + sL * sR * (UR[3] - UL[3])) * inv_S;
  out[4] = 0.0;
  out[5] = (sR * FL[5] - sL * FR[5] + sL * sR * (UR[5] - UL[5])) * inv_S;
}

/* computeHLLFluxY, HLL.hpp:177-218 */
void og_hll_y(const double D[6], const double U[6], double out[6]) {
  double FD[6], FU[6];
  convert_primitive_y(D, FD);
  convert_primitive_y(U, FU);
  double cD = og_fast_speed(D, D[5]);
  double cU = og_fast_speed(U, U[5]);
  double sD = dmin(D[2] - cD, U[2] - cU);
  double sU = dmax(D[2] + cD, U[2] + cU);
  if (sD >= 0) { memcpy(out, FD, sizeof FD); return; }
  if (sU <= 0) { memcpy(out, FU, sizeof FU); return; }
  double UD[6], UU[6];
  from_primitive(D, UD);
  from_primitive(U, UU);
  double inv_S = 1.0 / (sU - sD);
  out[0] = (sU * FD[0] - sD * FU[0] + sD * sU * (UU[0] - UD[0])) * inv_S;
  out[1] = (sU * FD[1] - sD * FU[1] + sD * sU * (UU[1] - UD[1])) * inv_S;
  out[2] = (sU * FD[2] - sD * FU[2] + sD * sU * (UU[2] - UD[2])) * inv_S;
  out[3] = (sU * FD[3] - sD * FU[3] + sD * sU * (UU[3] - UD[3])) * inv_S;
  out[4] = (sU * FD[4] - sD * FU[4] + sD * sU * (UU[4] - UD[4])) * inv_S;
  out[5] = 0.0;
}

/* ---------------- flux assembly and CT (include/solvers.hpp) ---------------- */

/* generate_cell_properties, solvers.hpp:216-251 (i, j physical) */
static void cell_properties(const og_grid* g, const og_ws* w, int i, int j,
                            double left[6], double right[6], double top[6], double bottom[6]) {
  int cur = IDX(g, i + g->ng, j + g->ng);
  int rgt = IDX(g, i + 1 + g->ng, j + g->ng);
  int topi = IDX(g, i + g->ng, j + 1 + g->ng);
  left[0] = w->rv[RV_RHO_XR][cur]; left[1] = w->rv[RV_VX_XR][cur]; left[2] = w->rv[RV_VY_XR][cur];
  left[3] = w->rv[RV_P_XR][cur]; left[4] = g->BXF[rgt]; left[5] = w->rv[RV_BY_XR][cur];
  right[0] = w->rv[RV_RHO_XL][rgt]; right[1] = w->rv[RV_VX_XL][rgt]; right[2] = w->rv[RV_VY_XL][rgt];
  right[3] = w->rv[RV_P_XL][rgt]; right[4] = g->BXF[rgt]; right[5] = w->rv[RV_BY_XL][rgt];
  top[0] = w->rv[RV_RHO_YL][topi]; top[1] = w->rv[RV_VX_YL][topi]; top[2] = w->rv[RV_VY_YL][topi];
  top[3] = w->rv[RV_P_YL][topi]; top[4] = w->rv[RV_BX_YL][topi]; top[5] = g->BYF[topi];
  bottom[0] = w->rv[RV_RHO_YR][cur]; bottom[1] = w->rv[RV_VX_YR][cur]; bottom[2] = w->rv[RV_VY_YR][cur];
  bottom[3] = w->rv[RV_P_YR][cur]; bottom[4] = w->rv[RV_BX_YR][cur]; bottom[5] = g->BYF[topi];
}

/* flux_population, solvers.hpp:348-391: cells [-1, n] */
void og_flux_population(og_grid* g, og_ws* w) {
  double L[6], R[6], T[6], B[6], fx[6], fy[6];
  for (int j = -1; j < g->ny + 1; j++) {
    for (int i = -1; i < g->nx + 1; i++) {
      int cur = IDX(g, i + g->ng, j + g->ng);
      cell_properties(g, w, i, j, L, R, T, B);
      og_hll_x(L, R, fx);
      og_hll_y(B, T, fy);
      w->fl[FL_RHO_H][cur] = fx[0]; w->fl[FL_MX_H][cur] = fx[1]; w->fl[FL_MY_H][cur] = fx[2];
      w->fl[FL_E_H][cur] = fx[3]; w->fl[FL_BX_H][cur] = fx[4]; w->fl[FL_BY_H][cur] = fx[5];
      w->fl[FL_RHO_V][cur] = fy[0]; w->fl[FL_MX_V][cur] = fy[1]; w->fl[FL_MY_V][cur] = fy[2];
      w->fl[FL_E_V][cur] = fy[3]; w->fl[FL_BX_V][cur] = fy[4]; w->fl[FL_BY_V][cur] = fy[5];
    }
  }
}

/* cell_corner_electric_fields, solvers.hpp:294-346: corners [-1, n].
 * The reference's unqualified abs() resolves to the double overload in its
 * real build (Eigen pulls in the libstdc++ <stdlib.h> wrapper); fabs here. */
void og_cell_corner_electric_fields(og_ws* w, og_grid* g) {
  for (int j = -1; j < g->ny + 1; j++) {
    for (int i = -1; i < g->nx + 1; i++) {
      int cur = IDX(g, i + g->ng, j + g->ng);
      int lft = IDX(g, i - 1 + g->ng, j + g->ng);
      int bot = IDX(g, i + g->ng, j - 1 + g->ng);
      int bl = IDX(g, i - 1 + g->ng, j - 1 + g->ng);
      double eN = -w->fl[FL_BY_H][lft];
      double eS = -w->fl[FL_BY_H][bl];
      double eE = w->fl[FL_BX_V][bot];
      double eW = w->fl[FL_BX_V][bl];
      double vxc = 0.25 * (g->VXF[cur] + g->VXF[lft] + g->VXF[bot] + g->VXF[bl]);
      double vyc = 0.25 * (g->VYF[cur] + g->VYF[lft] + g->VYF[bot] + g->VYF[bl]);
      double ezy, ezx;
      if (fabs(vyc) < 1e-12) ezy = 0.5 * (eS + eN);
      else ezy = (vyc > 0.0) ? eS : eN;
      if (fabs(vxc) < 1e-12) ezx = 0.5 * (eE + eW);
      else ezx = (vxc > 0.0) ? eW : eE;
      double e = 0.5 * (ezx + ezy);
      w->corner_e[cur] = e;
      g->EZF[cur] = e;
    }
  }
}

/* time_derivatives + get_partial_derivative, solvers.hpp:195-200, 253-292 */
void og_time_derivatives(og_ws* w, const og_grid* g, double* const d[6]) {
  const int nx = g->nx, ny = g->ny;
  const double dx = g->dx, dy = g->dy;
  for (int j = 0; j < ny; j++) {
    for (int i = 0; i < nx; i++) {
      int cur = IDX(g, i + 2, j + 2);
      int lft = IDX(g, i - 1 + 2, j + 2);
      int bot = IDX(g, i + 2, j - 1 + 2);
      d[0][cur] = -((w->fl[FL_RHO_H][cur] - w->fl[FL_RHO_H][lft]) / dx + (w->fl[FL_RHO_V][cur] - w->fl[FL_RHO_V][bot]) / dy);
      d[3][cur] = -((w->fl[FL_E_H][cur] - w->fl[FL_E_H][lft]) / dx + (w->fl[FL_E_V][cur] - w->fl[FL_E_V][bot]) / dy);
      d[1][cur] = -((w->fl[FL_MX_H][cur] - w->fl[FL_MX_H][lft]) / dx + (w->fl[FL_MX_V][cur] - w->fl[FL_MX_V][bot]) / dy);
      d[2][cur] = -((w->fl[FL_MY_H][cur] - w->fl[FL_MY_H][lft]) / dx + (w->fl[FL_MY_V][cur] - w->fl[FL_MY_V][bot]) / dy);
      /* periodic: the wrapped corner (solvers.hpp:278-290); outflow: the
       * boundary corner line, computed from the ghost-filled fluxes */
      int ir = g->bc_x ? i + 1 : (i + 1) % nx;
      int jt = g->bc_y ? j + 1 : (j + 1) % ny;
      int rgt2 = IDX(g, ir + 2, j + 2);
      int top2 = IDX(g, i + 2, jt + 2);
      d[4][cur] = -(g->EZF[top2] - g->EZF[cur]) / dy;
      d[5][cur] = (g->EZF[rgt2] - g->EZF[cur]) / dx;
    }
  }
}

/* explicit_rhs, solvers.hpp:393-402 */
void og_explicit_rhs(og_grid* g, og_ws* w, double* const d[6]) {
  og_fill_ghost(g);
  og_compute_primitives(g);
  og_muscl_all(g, w);
  og_flux_population(g, w);
  og_cell_corner_electric_fields(w, g);
  og_time_derivatives(w, g, d);
}

/* ---------------- implicit operator (solvers.hpp:404-489 + Eigen CG) ---------------- */

typedef struct { int row, col; double val; } triplet;

/* generate_A_matrix, solvers.hpp:404-443, followed by Eigen's setFromTriplets
 * (insert per row in triplet order, sum duplicates in order, transposed copy
 * which sorts inner indices) and DiagonalPreconditioner::compute (invdiag). */
void og_generate_A_matrix(const og_grid* g, og_ws* w, double alpha) {
  const int xs = g->nx, ys = g->ny, n = xs * ys;
  const double cx = alpha / (g->dx * g->dx);
  const double cy = alpha / (g->dy * g->dy);
  triplet* t = (triplet*)malloc((size_t)n * 5 * sizeof(triplet));
  size_t nt = 0;
  for (int j = 0; j < ys; j++) {
    for (int i = 0; i < xs; i++) {
      int c = j * xs + i;
      /* a neighbour past an outflow boundary is the corner itself (Neumann) */
      if (j == 0) t[nt++] = (triplet){c, g->bc_y ? c : (ys - 1) * xs + i, -cy};
      else t[nt++] = (triplet){c, c - xs, -cy};
      if (j == ys - 1) t[nt++] = (triplet){c, g->bc_y ? c : i, -cy};
      else t[nt++] = (triplet){c, c + xs, -cy};
      if (i == 0) t[nt++] = (triplet){c, g->bc_x ? c : c + xs - 1, -cx};
      else t[nt++] = (triplet){c, c - 1, -cx};
      if (i == xs - 1) t[nt++] = (triplet){c, g->bc_x ? c : c - xs + 1, -cx};
      else t[nt++] = (triplet){c, c + 1, -cx};
      t[nt++] = (triplet){c, c, 1 + 2 * (cx + cy)};
    }
  }
  /* pass 1-2: the transposed (row-major) matrix, inner index = col, insertion order */
  int* tptr = (int*)calloc((size_t)n + 1, sizeof(int));
  for (size_t k = 0; k < nt; ++k) tptr[t[k].row + 1]++;
  for (int r = 0; r < n; ++r) tptr[r + 1] += tptr[r];
  int* tcol = (int*)malloc(nt * sizeof(int));
  double* tval = (double*)malloc(nt * sizeof(double));
  int* fillp = (int*)malloc((size_t)n * sizeof(int));
  memcpy(fillp, tptr, (size_t)n * sizeof(int));
  for (size_t k = 0; k < nt; ++k) { int pos = fillp[t[k].row]++; tcol[pos] = t[k].col; tval[pos] = t[k].val; }
  free(t);
  /* pass 3: collapse duplicates within each row, in order of appearance */
  int* cnt = (int*)calloc((size_t)n, sizeof(int));
  for (int r = 0; r < n; ++r) {
    int beg = tptr[r], end = tptr[r + 1], m = 0;
    for (int k = beg; k < end; ++k) {
      int found = -1;
      for (int q = beg; q < beg + m; ++q) if (tcol[q] == tcol[k]) { found = q; break; }
      if (found >= 0) tval[found] += tval[k];
      else { tcol[beg + m] = tcol[k]; tval[beg + m] = tval[k]; ++m; }
    }
    cnt[r] = m;
  }
  /* pass 4: transposed copy into the column-major matrix (inner = row, sorted).
   * Column c of that matrix is what CG's transposed view reads as row c. */
  int* ptr = (int*)calloc((size_t)n + 1, sizeof(int));
  for (int r = 0; r < n; ++r) for (int k = 0; k < cnt[r]; ++k) ptr[tcol[tptr[r] + k] + 1]++;
  for (int c = 0; c < n; ++c) ptr[c + 1] += ptr[c];
  int nnz = ptr[n];
  int* col = (int*)malloc((size_t)nnz * sizeof(int));
  double* val = (double*)malloc((size_t)nnz * sizeof(double));
  memcpy(fillp, ptr, (size_t)n * sizeof(int));
  for (int r = 0; r < n; ++r) {
    for (int k = 0; k < cnt[r]; ++k) {
      int c = tcol[tptr[r] + k];
      int pos = fillp[c]++;
      col[pos] = r; val[pos] = tval[tptr[r] + k];
    }
  }
  free(tptr); free(tcol); free(tval); free(fillp); free(cnt);
  /* matrix_store.A_matrix = next_time_step_matrix (a copy), solver.compute(A) */
  free(w->a_ptr); free(w->a_col); free(w->a_val);
  w->a_ptr = ptr; w->a_col = col; w->a_val = val;
  for (int c = 0; c < n; ++c) {
    double inv = 1.0;
    for (int k = ptr[c]; k < ptr[c + 1]; ++k) if (col[k] == c) inv = 1.0 / val[k];
    w->invdiag[c] = inv;
  }
}

static void spmv(const og_ws* w, const double* x, double* y) {
  const int n = w->nphys;
  for (int i = 0; i < n; ++i) {
    double s = 0;
    for (int k = w->a_ptr[i]; k < w->a_ptr[i + 1]; ++k) s += w->a_val[k] * x[w->a_col[k]];
    y[i] = s;
  }
}

static double dot(const double* a, const double* b, int n) {
  double s = 0;
  for (int i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

void og_ws_set_cg(og_ws* w, double tol, long long max_iter) { w->cg_tol = tol; w->cg_maxit = max_iter; }

/* Eigen::ConjugateGradient<SparseMatrix<double>, Lower|Upper>::solve(b) with the
 * default DiagonalPreconditioner and a zero initial guess (Eigen's
 * internal::conjugate_gradient, restated from its published algorithm). */
void og_cg_solve(og_ws* w, const double* b, double* x, int* iters, int* nspmv, double* err) {
  const int n = w->nphys;
  long long maxIters = (w->cg_maxit < 0) ? 2LL * n : w->cg_maxit;
  double tol = w->cg_tol;
  double* r = w->r; double* p = w->p; double* z = w->z; double* tmp = w->tmp;
  int spmvs = 0;
  for (int i = 0; i < n; ++i) x[i] = 0.0;
  for (int i = 0; i < n; ++i) r[i] = b[i] - 0.0;
  double rhsNorm2 = dot(b, b, n);
  if (rhsNorm2 == 0) {
    for (int i = 0; i < n; ++i) x[i] = 0.0;
    *iters = 0; *nspmv = 0; *err = 0; return;
  }
  double considerAsZero = DBL_MIN;
  double threshold = dmax(tol * tol * rhsNorm2, considerAsZero);
  double residualNorm2 = dot(r, r, n);
  if (residualNorm2 < threshold) {
    *iters = 0; *nspmv = 0; *err = sqrt(residualNorm2 / rhsNorm2); return;
  }
  for (int i = 0; i < n; ++i) p[i] = w->invdiag[i] * r[i];
  double absNew = dot(r, p, n);
  long long it = 0;
  while (it < maxIters) {
    spmv(w, p, tmp); ++spmvs;
    double alpha = absNew / dot(p, tmp, n);
    for (int i = 0; i < n; ++i) x[i] += alpha * p[i];
    for (int i = 0; i < n; ++i) r[i] -= alpha * tmp[i];
    residualNorm2 = dot(r, r, n);
    if (residualNorm2 < threshold) break;
    for (int i = 0; i < n; ++i) z[i] = w->invdiag[i] * r[i];
    double absOld = absNew;
    absNew = dot(r, z, n);
    double beta = absNew / absOld;
    for (int i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    it++;
  }
  *err = sqrt(residualNorm2 / rhsNorm2);
  *iters = (int)it;
  *nspmv = spmvs;
}

/* resistive_B_term, solvers.hpp:445-489 */
void og_resistive_term(const og_grid* g, og_ws* w, int implicit, double* bx_rate, double* by_rate,
                       int* iters, int* nspmv, double* err) {
  const int ny = g->ny, nx = g->nx;
  for (int j = 0; j < ny; j++) {
    for (int i = 0; i < nx; i++) {
      int cur = IDX(g, i + 2, j + 2);
      int lft = IDX(g, i - 1 + 2, j + 2);
      int bot = IDX(g, i + 2, j - 1 + 2);
      double dBy_dx = (g->BYF[cur] - g->BYF[lft]) / g->dx;
      double dBx_dy = (g->BXF[cur] - g->BXF[bot]) / g->dy;
      w->ez_compact[j * nx + i] = g->eta * (dBy_dx - dBx_dy);
    }
  }
  const double* ez = w->ez_compact;
  int it = 0, sp = 0; double e = 0;
  if (implicit) {
    og_cg_solve(w, w->ez_compact, w->xs, &it, &sp, &e);
    ez = w->xs;
  }
  if (iters) *iters = it;
  if (nspmv) *nspmv = sp;
  if (err) *err = e;
  for (int j = 0; j < ny; j++) {
    for (int i = 0; i < nx; i++) {
      int ir = g->bc_x ? clamp_index(i + 1, nx) : (i + 1) % nx;
      int jt = g->bc_y ? clamp_index(j + 1, ny) : (j + 1) % ny;
      double ec = ez[j * nx + i], et = ez[jt * nx + i], er = ez[j * nx + ir];
      int gidx = IDX(g, i + 2, j + 2);
      bx_rate[gidx] = -(et - ec) / g->dy;
      by_rate[gidx] = (er - ec) / g->dx;
    }
  }
}

/* ---------------- IMEX ARS(2,2,2) (solvers.hpp:39-66, 202-214, 491-613) ---------------- */

/* get_intermediary_grid, solvers.hpp:202-214 */
static void axpy_phys(const og_grid* g, double* prop, const double* change, double s) {
  for (int j = 0; j < g->ny; j++)
    for (int i = 0; i < g->nx; i++) {
      int id = IDX(g, i + 2, j + 2);
      prop[id] += change[id] * s;
    }
}

void og_imex(og_grid* gi, og_ws* w, double dt, og_step_stats* st) {
  /* IMEX_Butcher_Tableus, solvers.hpp:41-64 */
  const double gam = (2.0 - sqrt(2.0)) / 2.0;
  const double del = 1.0 - (1.0 / (2.0 * gam));
  const double A_ex3 = gam, A_ex6 = del, A_ex7 = 1.0 - del;
  const double A_im3 = 0.0, A_im4 = gam, A_im6 = 0.0, A_im7 = 1.0 - gam;
  const double b_ex0 = del, b_ex1 = 1.0 - del, b_im1 = 1.0 - gam, b_im2 = gam;
  og_step_stats dummy;
  if (!st) st = &dummy;
  memset(st, 0, sizeof *st);

  og_grid* initial = grid_clone(gi);                         /* :493 */
  double* f1[6] = {w->f1[0], w->f1[1], w->f1[2], w->f1[3], w->f1[4], w->f1[5]};
  double* f2[6] = {w->f2[0], w->f2[1], w->f2[2], w->f2[3], w->f2[4], w->f2[5]};
  og_explicit_rhs(gi, w, f1);                                /* :495-496 */
  og_resistive_term(gi, w, 0, w->g1[0], w->g1[1], NULL, NULL, NULL);   /* :499 */
  og_grid* inter = w->inter; og_grid* whole = w->whole;
  og_grid_copy(inter, initial);                              /* :501 */
  og_grid_copy(whole, initial);                              /* :502 */
  const double diag_coeff = dt * A_im4;                      /* :503 */

  const double ce1 = dt * A_ex3, ci1 = dt * A_im3;           /* :509-510 */
  axpy_phys(gi, inter->RHO, f1[0], ce1);
  axpy_phys(gi, inter->MXF, f1[1], ce1);
  axpy_phys(gi, inter->MYF, f1[2], ce1);
  axpy_phys(gi, inter->EF, f1[3], ce1);
  axpy_phys(gi, inter->BXF, f1[4], ce1);
  axpy_phys(gi, inter->BYF, f1[5], ce1);
  axpy_phys(gi, inter->BXF, w->g1[0], ci1);
  axpy_phys(gi, inter->BYF, w->g1[1], ci1);
  og_fill_ghost(inter);
  og_compute_primitives(inter);
  og_resistive_term(inter, w, 1, w->solved1[0], w->solved1[1],
                    &st->cg_iters[0], &st->cg_spmv[0], &st->cg_error[0]);   /* :526 */
  axpy_phys(gi, inter->BXF, w->solved1[0], diag_coeff);
  axpy_phys(gi, inter->BYF, w->solved1[1], diag_coeff);
  og_fill_ghost(inter);
  og_compute_primitives(inter);
  og_explicit_rhs(inter, w, f2);                             /* :534-535 */
  og_resistive_term(inter, w, 0, w->g2[0], w->g2[1], NULL, NULL, NULL);   /* :537 */

  const double ce2 = dt * A_ex6, ce3 = dt * A_ex7, ci2 = dt * A_im6, ci3 = dt * A_im7;   /* :543-546 */
  axpy_phys(gi, whole->RHO, f1[0], ce2);
  axpy_phys(gi, whole->MXF, f1[1], ce2);
  axpy_phys(gi, whole->MYF, f1[2], ce2);
  axpy_phys(gi, whole->EF, f1[3], ce2);
  axpy_phys(gi, whole->BXF, f1[4], ce2);
  axpy_phys(gi, whole->BYF, f1[5], ce2);
  axpy_phys(gi, whole->RHO, f2[0], ce3);
  axpy_phys(gi, whole->MXF, f2[1], ce3);
  axpy_phys(gi, whole->MYF, f2[2], ce3);
  axpy_phys(gi, whole->EF, f2[3], ce3);
  axpy_phys(gi, whole->BXF, f2[4], ce3);
  axpy_phys(gi, whole->BYF, f2[5], ce3);
  axpy_phys(gi, whole->BXF, w->g1[0], ci2);
  axpy_phys(gi, whole->BYF, w->g1[1], ci2);
  axpy_phys(gi, whole->BXF, w->g2[0], ci3);
  axpy_phys(gi, whole->BYF, w->g2[1], ci3);
  og_fill_ghost(whole);
  og_compute_primitives(whole);
  og_resistive_term(whole, w, 1, w->solved2[0], w->solved2[1],
                    &st->cg_iters[1], &st->cg_spmv[1], &st->cg_error[1]);   /* :571 */
  axpy_phys(gi, whole->BXF, w->solved2[0], diag_coeff);
  axpy_phys(gi, whole->BYF, w->solved2[1], diag_coeff);
  og_fill_ghost(whole);
  og_compute_primitives(whole);
  og_resistive_term(whole, w, 0, w->g3[0], w->g3[1], NULL, NULL, NULL);   /* :579 */

  const double f1c = dt * b_ex0, f2c = dt * b_ex1, g2c = dt * b_im1, g3c = dt * b_im2;   /* :585-588 */
  axpy_phys(gi, gi->RHO, f1[0], f1c);
  axpy_phys(gi, gi->MXF, f1[1], f1c);
  axpy_phys(gi, gi->MYF, f1[2], f1c);
  axpy_phys(gi, gi->EF, f1[3], f1c);
  axpy_phys(gi, gi->BXF, f1[4], f1c);
  axpy_phys(gi, gi->BYF, f1[5], f1c);
  axpy_phys(gi, gi->RHO, f2[0], f2c);
  axpy_phys(gi, gi->MXF, f2[1], f2c);
  axpy_phys(gi, gi->MYF, f2[2], f2c);
  axpy_phys(gi, gi->EF, f2[3], f2c);
  axpy_phys(gi, gi->BXF, f2[4], f2c);
  axpy_phys(gi, gi->BYF, f2[5], f2c);
  axpy_phys(gi, gi->BXF, w->g2[0], g2c);
  axpy_phys(gi, gi->BYF, w->g2[1], g2c);
  axpy_phys(gi, gi->BXF, w->g3[0], g3c);
  axpy_phys(gi, gi->BYF, w->g3[1], g3c);
  og_fill_ghost(gi);
  og_compute_primitives(gi);
  og_grid_destroy(initial);
}

/* ---------------- diagnostics (include/diagnostics.hpp) ---------------- */

/* calculate_current_div_B, diagnostics.hpp:15-42 */
void og_divb(const og_grid* g, double* rms_out, double* max_out) {
  double mx = 0.0, rms = 0.0;
  for (int j = 0; j < g->ny; j++) {
    for (int i = 0; i < g->nx; i++) {
      int cur = IDX(g, i + 2, j + 2), rgt = IDX(g, i + 3, j + 2), top = IDX(g, i + 2, j + 3);
      double xd = (g->BXF[rgt] - g->BXF[cur]) / g->dx;
      double yd = (g->BYF[top] - g->BYF[cur]) / g->dy;
      double t = xd + yd;
      rms += t * t;
      if (fabs(t) > mx) mx = fabs(t);
    }
  }
  rms = sqrt(rms / (g->nx * g->ny));
  *rms_out = rms; *max_out = mx;
}

/* CFL_condition_checker, diagnostics.hpp:44-88.  The reference hard-codes an
 * effective Courant number 0.5 * 0.9 = 0.45; here it is the argument `cfl`,
 * applied as (cfl / 0.9) * h / max_speed before the 0.9 safety factor, which
 * reproduces the reference exactly for cfl = 0.45 (0.45 / 0.9 == 0.5 in binary64). */
double og_cfl(const og_grid* g, double cfl) {
  const double dx = g->dx, dy = g->dy, adiabatic_index = 5.0 / 3.0;
  const double courant = cfl / 0.9;
  double dt_res = INFINITY;
  if (g->eta != 0.0) dt_res = dmin(dx, dy) * dmin(dx, dy) / (4.0 * g->eta);
  double maxs = 1e-8;
  for (int j = 0; j < g->ny; j++) {
    for (int i = 0; i < g->nx; i++) {
      int id = IDX(g, i + 2, j + 2);
      double rho = dmax(g->RHO[id], 1e-8);
      double p = dmax(g->PF[id], 1e-8);
      double vx = g->VXF[id], vy = g->VYF[id], bx = g->BXC[id], by = g->BYC[id];
      double ss = adiabatic_index * p / rho;
      double as = (bx * bx + by * by) / rho;
      double ms = sqrt(ss + as);
      double fs = sqrt(vx * vx + vy * vy);
      double cur = ms + fs;
      if (cur > maxs) maxs = cur;
    }
  }
  double dt_hyp = courant * dmin(dx, dy) / maxs;
  double lim = dmin(dt_hyp, dt_res);
  return lim * 0.9;
}

/* ---------------- initial conditions (include/benchmarks.hpp) ---------------- */

static void store_state(og_grid* g, int idx, const double w[6], int b_from_u) {
  double u[6];
  from_primitive(w, u);
  g->RHO[idx] = u[0]; g->MXF[idx] = u[1]; g->MYF[idx] = u[2]; g->EF[idx] = u[3];
  g->BXF[idx] = b_from_u ? u[4] : w[4];
  g->BYF[idx] = b_from_u ? u[5] : w[5];
  g->PF[idx] = w[3]; g->VXF[idx] = w[1]; g->VYF[idx] = w[2];
}

/* orszag_tang_initial_conditions, benchmarks.hpp:15-76 (including the
 * indexC(nxt, j) aliasing of the potential array, which only touches ghosts) */
void og_ic_orszag_tang(og_grid* g, og_ws* w) {
  const int nxt = g->nxt, nyt = g->nyt, ng = g->ng;
  const double dx = g->dx, dy = g->dy;
  double* Az = w->potential;
  const double rho = 25.0 / 9.0, p = 5.0 / 3.0;
  for (int j = 0; j <= nyt; ++j)
    for (int i = 0; i <= nxt; ++i) {
      double xc = (i - ng) * dx, yc = (j - ng) * dy;
      Az[IDX(g, i, j)] = 0.5 * cos(2 * xc) + cos(yc);
    }
  for (int j = 0; j < nyt; ++j)
    for (int i = 0; i < nxt; ++i) {
      int idx = IDX(g, i, j), top = IDX(g, i, j + 1), rgt = IDX(g, i + 1, j);
      double x = (i - ng + 0.5) * dx, y = (j - ng + 0.5) * dy;
      double wv[6] = {rho, -sin(y), sin(x), p, (Az[top] - Az[idx]) / dy, -(Az[rgt] - Az[idx]) / dx};
      store_state(g, idx, wv, 1);
    }
}

/* alfven_wave_test_initial_values, benchmarks.hpp:78-126 */
void og_ic_alfven(og_grid* g) {
  const double pi = PI_REF;
  const double W = g->nx * g->dx;
  const double kx = 2.0 * pi / W;
  for (int j = 0; j < g->nyt; ++j)
    for (int i = 0; i < g->nxt; ++i) {
      double x = (i - g->ng + 0.5) * g->dx;
      double by = 1e-3 * sin(kx * x);
      double vy = -by / sqrt(1.0);
      double wv[6] = {1.0, 0.0, vy, 1.0, 1.0, by};
      store_state(g, IDX(g, i, j), wv, 1);
    }
}

/* gaussian_density_pulse_test_initial_values, benchmarks.hpp:128-176 */
void og_ic_gaussian(og_grid* g) {
  const double x_mid = 0.5 * g->nx * g->dx;
  const double y_mid = 0.5 * g->nx * g->dy;
  const double radius_sq = 2.0 * 2.0;
  for (int j = 0; j < g->nyt; ++j)
    for (int i = 0; i < g->nxt; ++i) {
      double x = (i - g->ng + 0.5) * g->dx, y = (j - g->ng + 0.5) * g->dy;
      double d2 = (x - x_mid) * (x - x_mid) + (y - y_mid) * (y - y_mid);
      double rho = 1.0 + exp(-5 * d2 / radius_sq);
      double wv[6] = {rho, 1.0, 1.0, 1.0, 0.0, 0.0};
      store_state(g, IDX(g, i, j), wv, 1);
    }
}

/* acoustic_wave_test_initial_values, benchmarks.hpp:178-231 */
void og_ic_acoustic(og_grid* g) {
  const double pi = PI_REF, gam = 5.0 / 3.0;
  const double W = g->dx * g->nx;
  const double p_bg = 1.0 / gam, rho_bg = 1.0;
  const double cs = sqrt(gam * p_bg / rho_bg);
  for (int j = 0; j < g->nyt; ++j)
    for (int i = 0; i < g->nxt; ++i) {
      double x = (i - g->ng) * g->dx;
      double xs = 1e-4 * sin(4 * pi * x / W), ys = 0.0;
      double rho = rho_bg + xs + ys;
      double vx = (cs / rho_bg) * xs, vy = (cs / rho_bg) * ys;
      double p = p_bg + (cs * cs) * (xs + ys);
      double wv[6] = {rho, vx, vy, p, 0.0, 0.0};
      store_state(g, IDX(g, i, j), wv, 1);
    }
}

/* sausage_instability_initial_conditions, benchmarks.hpp:233-293 */
void og_ic_sausage(og_grid* g) {
  const double pi = PI_REF;
  const double H = g->ny * g->dy;
  const double k = 2.0 * pi / H, radius_sq = 0.04;
  const double x_mid = 0.5 * g->nxt * g->dx;
  for (int j = 0; j < g->nyt; ++j)
    for (int i = 0; i < g->nxt; ++i) {
      double x = (i - g->ng + 0.5) * g->dx, y = (j - g->ng + 0.5) * g->dy;
      double r = x - x_mid;
      double pert = 1.0 - 0.15 * cos(k * y);
      double rho = 1.0 * exp(-(r * r) / (radius_sq * pert)) + 0.1;
      double p = 0.1 * exp(-(r * r) / (radius_sq * pert)) + 0.1;
      rho = dmax(rho, 1e-4);
      p = dmax(p, 1e-4);
      double by = 5.0 * (r / sqrt(radius_sq)) * exp(-(r * r) / radius_sq);
      double wv[6] = {rho, 0.0, 0.0, p, 0.0, by};
      store_state(g, IDX(g, i, j), wv, 0);
    }
}

/* kink_instability_initial_conditions, benchmarks.hpp:295-350 */
void og_ic_kink(og_grid* g) {
  const double pi = PI_REF;
  const double k = 2.0 * pi, radius_sq = 0.04;
  const double x_mid = 0.5 * g->nxt * g->dx;
  for (int j = 0; j < g->nyt; ++j)
    for (int i = 0; i < g->nxt; ++i) {
      double x = (i - g->ng + 0.5) * g->dx, y = (j - g->ng + 0.5) * g->dy;
      double r = x - x_mid;
      double env = exp(-(r * r) / radius_sq);
      double wv[6] = {1.0 * env + 0.1, 0.05 * sin(k * y) * env, 0.0, 1.0 * env + 0.1, 0.0, 1.0 * env};
      store_state(g, IDX(g, i, j), wv, 0);
    }
}

/* decaying_magnetic_sine_wave_initial_conditions, benchmarks.hpp:352-417 */
void og_ic_decaying_sine(og_grid* g, og_ws* w) {
  const double pi = PI_REF;
  const double W = g->nx * g->dx, H = g->ny * g->dy;
  const double kx = 2 * pi / W, ky = 2 * pi / H;
  double* Az = w->potential;
  for (int j = 0; j <= g->nyt; j++)
    for (int i = 0; i <= g->nxt; i++) {
      double xl = (i - g->ng) * g->dx, yb = (j - g->ng) * g->dy;
      Az[IDX(g, i, j)] = (1.0 / kx) * cos(kx * xl) * sin(ky * yb);
    }
  for (int j = 0; j < g->nyt; j++)
    for (int i = 0; i < g->nxt; i++) {
      int bl = IDX(g, i, j), br = IDX(g, i + 1, j), tl = IDX(g, i, j + 1);
      double bx = (Az[tl] - Az[bl]) / g->dy;
      double by = -(Az[br] - Az[bl]) / g->dx;
      double wv[6] = {1.0, 0.0, 0.0, 1.0, bx, by};
      store_state(g, bl, wv, 0);
    }
}

/* ---------------- workspace ---------------- */

og_ws* og_ws_create(int nx, int ny) {
  og_ws* w = (og_ws*)calloc(1, sizeof(og_ws));
  w->nx = nx; w->ny = ny; w->nxt = nx + 4; w->nyt = ny + 4;
  w->ncell = w->nxt * w->nyt; w->nphys = nx * ny;
  size_t n = (size_t)w->ncell;
  for (int k = 0; k < RV_COUNT; ++k) w->rv[k] = dalloc(n);
  for (int k = 0; k < FL_COUNT; ++k) w->fl[k] = dalloc(n);
  for (int k = 0; k < 6; ++k) { w->f1[k] = dalloc(n); w->f2[k] = dalloc(n); }
  for (int k = 0; k < 2; ++k) {
    w->g1[k] = dalloc(n); w->g2[k] = dalloc(n); w->g3[k] = dalloc(n);
    w->solved1[k] = dalloc(n); w->solved2[k] = dalloc(n);
  }
  w->corner_e = dalloc(n);
  w->ez_compact = dalloc((size_t)w->nphys);
  w->potential = dalloc((size_t)(w->nxt + 1) * (w->nyt + 1));
  w->invdiag = dalloc((size_t)w->nphys);
  w->r = dalloc((size_t)w->nphys); w->p = dalloc((size_t)w->nphys);
  w->z = dalloc((size_t)w->nphys); w->tmp = dalloc((size_t)w->nphys);
  w->xs = dalloc((size_t)w->nphys);
  w->cg_tol = DBL_EPSILON; w->cg_maxit = -1;
  w->inter = og_grid_create(nx, ny, 0, 0, 0);
  w->whole = og_grid_create(nx, ny, 0, 0, 0);
  return w;
}

void og_ws_destroy(og_ws* w) {
  if (!w) return;
  for (int k = 0; k < RV_COUNT; ++k) free(w->rv[k]);
  for (int k = 0; k < FL_COUNT; ++k) free(w->fl[k]);
  for (int k = 0; k < 6; ++k) { free(w->f1[k]); free(w->f2[k]); }
  for (int k = 0; k < 2; ++k) { free(w->g1[k]); free(w->g2[k]); free(w->g3[k]); free(w->solved1[k]); free(w->solved2[k]); }
  free(w->corner_e); free(w->ez_compact); free(w->potential); free(w->invdiag);
  free(w->r); free(w->p); free(w->z); free(w->tmp); free(w->xs);
  free(w->a_ptr); free(w->a_col); free(w->a_val);
  og_grid_destroy(w->inter); og_grid_destroy(w->whole);
  free(w);
}

double* og_ws_rv(og_ws* w, int which) { return (which >= 0 && which < RV_COUNT) ? w->rv[which] : NULL; }
double* og_ws_flux(og_ws* w, int which) { return (which >= 0 && which < FL_COUNT) ? w->fl[which] : NULL; }

/* ---------------- driver (src/integrator_main.cpp:52-112) ---------------- */

int og_run(og_grid* g, og_ws* w, double t_end, double dt_override, double cfl,
           int max_steps, int snapshot_every, double* step_dt, int* iters, double* t_out) {
  const double A_im4 = (2.0 - sqrt(2.0)) / 2.0;
  og_fill_ghost(g);                                          /* :52-53 */
  og_compute_primitives(g);
  double opt = og_cfl(g, cfl);                               /* :55 */
  double dt0 = (dt_override == 0.0) ? opt : dt_override;
  double mc = dt0 * g->eta;
  og_generate_A_matrix(g, w, mc * A_im4);                    /* :58-60 */
  og_compute_primitives(g);                                  /* :67 */
  double prev = dt0, t = 0.0;
  int frame = 0, steps = 0;
  while (t < t_end && (max_steps < 0 || steps < max_steps)) {   /* :76 */
    double copt = og_cfl(g, cfl);
    double dt = (dt_override == 0.0) ? copt : dt_override;
    if (t + dt > t_end) dt = t_end - t;                      /* :83 */
    frame++;
    if (g->eta > 0.0) {
      if (fabs(dt - prev) > 1e-10 * dt) {                    /* :88 */
        double m = dt * g->eta;
        og_generate_A_matrix(g, w, m * A_im4);
      }
    }
    prev = dt;
    og_step_stats st;
    og_imex(g, w, dt, &st);                                  /* :99 */
    t += dt;
    if (step_dt) step_dt[steps] = dt;
    if (iters) { iters[2 * steps] = st.cg_iters[0]; iters[2 * steps + 1] = st.cg_iters[1]; }
    steps++;
    if (snapshot_every > 0 && frame % snapshot_every == 0) {   /* :102-111 */
      og_fill_ghost(g);
      og_compute_primitives(g);
    }
  }
  if (t_out) *t_out = t;
  return steps;
}
