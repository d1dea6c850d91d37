/*
 * mhd_oracle.h -- CPU restatement of the reference MHD step (TEST INFRASTRUCTURE).
 *
 * This library is the parity checker for the HIP path, nothing more.  Only the
 * test suite, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it.  The product (magnetohydrodynamics_simulator_amd) never links or calls it.
 *
 * It restates, in plain C, the algorithm of the reference's header-only solver
 * (paths relative to the reference repository):
 *   include/Grid.hpp          padded SoA grid, periodic ghost fill, primitives
 *   include/MUSCL.hpp         minmod MUSCL reconstruction
 *   include/HLL.hpp           HLL x/y Riemann fluxes and fast magnetosonic speed
 *   include/solvers.hpp       flux assembly, upwind CT, resistive term, IMEX
 *   include/diagnostics.hpp   CFL time step and div(B) diagnostics
 *   include/benchmarks.hpp    initial conditions
 *   src/integrator_main.cpp   driver loop (dt, clamp, operator rebuild)
 * Eigen's ConjugateGradient (third-party, version unpinned by the reference's
 * CMakeLists.txt:7, absent from this machine) is restated from its published
 * algorithm: Jacobi preconditioner, zero initial guess, stop when
 * |r|^2 < max(tol^2 |b|^2, DBL_MIN), maxIter = 2n, iteration count = number of
 * completed non-final iterations.
 *
 * Pinning (what each function is checked against) is listed in the test files
 * under tests/ and in the header comment of mhd_oracle.c.
 */
#ifndef MHD_ORACLE_H
#define MHD_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* field ids for og_field(): the reference Grid members (include/Grid.hpp:19-26) */
enum {
  OG_RHO = 0, OG_MX = 1, OG_MY = 2, OG_E = 3, OG_BXF = 4, OG_BYF = 5,
  OG_VX = 6, OG_VY = 7, OG_P = 8, OG_BXC = 9, OG_BYC = 10, OG_EZ = 11,
  OG_NFIELDS = 12
};

typedef struct og_grid og_grid;
typedef struct og_ws og_ws;

typedef struct og_step_stats {
  int cg_iters[2];      /* Eigen iteration count of the two implicit solves */
  int cg_spmv[2];       /* matrix-vector products actually performed */
  double cg_error[2];   /* sqrt(|r|^2 / |b|^2) at exit (Eigen's tol_error) */
} og_step_stats;

/* grid (include/Grid.hpp:6-62) */
og_grid* og_grid_create(int nx, int ny, double dx, double dy, double eta);
void og_grid_destroy(og_grid* g);
double* og_field(og_grid* g, int which);
void og_grid_copy(og_grid* dst, const og_grid* src);
void og_grid_set_eta(og_grid* g, double eta);
/* 0 periodic (the reference), 1 outflow (zero-gradient ghost rule) per axis */
void og_grid_set_bc(og_grid* g, int bc_x, int bc_y);
void og_fill_ghost(og_grid* g);
void og_fill_ghost_periodic(og_grid* g);
void og_compute_cell_b(og_grid* g);
void og_compute_primitives(og_grid* g);

/* workspace: integrator_reserved_memory + ReconstructedValues + driver grids */
og_ws* og_ws_create(int nx, int ny);
void og_ws_destroy(og_ws* w);
void og_ws_set_cg(og_ws* w, double tol, long long max_iter);
double* og_ws_rv(og_ws* w, int which);       /* 20 MUSCL face arrays, MUSCL.hpp:52-58 order */
double* og_ws_flux(og_ws* w, int which);     /* 12 flux arrays, solvers.hpp:120-141 order */

/* scalar kernels */
double og_minmod(double a, double b);
void og_muscl_one(const double* F, const og_grid* g, int i, int j, int dir, double* L, double* R);
double og_fast_speed(const double prim[6], double bn);
void og_hll_x(const double L[6], const double R[6], double out[6]);
void og_hll_y(const double D[6], const double U[6], double out[6]);

/* spatial operator passes */
void og_muscl_all(const og_grid* g, og_ws* w);
void og_flux_population(og_grid* g, og_ws* w);
void og_cell_corner_electric_fields(og_ws* w, og_grid* g);
void og_time_derivatives(og_ws* w, const og_grid* g, double* const d[6]);
/* d = {rho_dot, mx_dot, my_dot, E_dot, bxf_dot, byf_dot}, padded arrays */
void og_explicit_rhs(og_grid* g, og_ws* w, double* const d[6]);

/* implicit resistive operator */
void og_generate_A_matrix(const og_grid* g, og_ws* w, double alpha);
void og_cg_solve(og_ws* w, const double* b, double* x, int* iters, int* spmv, double* err);
void og_resistive_term(const og_grid* g, og_ws* w, int implicit, double* bx_rate, double* by_rate,
                       int* iters, int* spmv, double* err);

/* one ARS(2,2,2) step, include/solvers.hpp:491-613 */
void og_imex(og_grid* g, og_ws* w, double dt, og_step_stats* st);

/* diagnostics */
double og_cfl(const og_grid* g, double cfl);
void og_divb(const og_grid* g, double* rms, double* maxabs);

/* initial conditions (include/benchmarks.hpp) */
void og_ic_orszag_tang(og_grid* g, og_ws* w);
void og_ic_alfven(og_grid* g);
void og_ic_gaussian(og_grid* g);
void og_ic_acoustic(og_grid* g);
void og_ic_sausage(og_grid* g);
void og_ic_kink(og_grid* g);
void og_ic_decaying_sine(og_grid* g, og_ws* w);

/* driver loop, src/integrator_main.cpp:52-112.  Returns number of steps taken.
 * dt_override == 0 selects the CFL time step.  snapshot_every > 0 repeats the
 * driver's ghost fill + primitives at snapshot frames (no file is written).
 * If step_dt != NULL it receives each step's dt (capacity max_steps);
 * if iters != NULL it receives 2 CG iteration counts per step. */
int og_run(og_grid* g, og_ws* w, double t_end, double dt_override, double cfl,
           int max_steps, int snapshot_every, double* step_dt, int* iters, double* t_out);

#ifdef __cplusplus
}
#endif
#endif
