/*
 * mhd_mi355x.h -- C ABI of the MI355X (gfx950) 2D resistive-MHD step.
 *
 * Drop-in replacement for the step interface of the reference solver
 * (RikibabManos/Magnetohydrodynamics_Simulator; citations relative to its
 * repository).  The reference exposes header-only C++ free functions that
 * mutate caller-owned std::vector<double> fields of a padded `Grid`
 * (include/Grid.hpp:6-62).  This ABI keeps that padded row-major layout at the
 * boundary -- every host array is (ny + 4) x (nx + 4) doubles, index
 * j * (nx + 4) + i, two ghost layers (Grid.hpp:29-58) -- so a host Grid's
 * .data() pointers pass straight through.
 *
 * Conventions
 *   - every entry point returns 0 (MHD_OK) or a positive MHD_E* code;
 *     mhd_last_error(ctx) describes the last failure on that context
 *     (mhd_last_error(NULL) the last failed mhd_create on this thread);
 *   - the library owns all device memory and never keeps a host pointer
 *     after a call returns;
 *   - a context is used from one host thread at a time (the reference is
 *     single-threaded and non-reentrant as well);
 *   - all arithmetic is IEEE binary64 on the GPU; there is no CPU fallback:
 *     without a usable gfx950 device mhd_create fails with MHD_EHIP.
 */
#ifndef MHD_MI355X_H
#define MHD_MI355X_H

#ifdef __cplusplus
extern "C" {
#endif

/* Entry points are exported; everything else in the library is internal. */
#if defined(__GNUC__) || defined(__clang__)
#define MHD_API __attribute__((visibility("default")))
#else
#define MHD_API
#endif

#define MHD_OK 0
#define MHD_EINVAL 1       /* bad argument (size, pointer, option) */
#define MHD_EHIP 2         /* HIP runtime error or no device */
#define MHD_ENONFINITE 3   /* a non-finite value was produced (CFL speed) */
#define MHD_ECG 4          /* CG reached max iterations (only if opts.cg_fail_is_error) */
#define MHD_ENOMEM 5       /* device allocation failed */

typedef struct mhd_ctx mhd_ctx;

/* Solver options.  Defaults reproduce the reference's hard-coded constants. */
typedef struct mhd_opts {
  double gamma;            /* adiabatic index; reference: 5/3 (Grid.hpp:104, HLL.hpp:31) -- informational,
                              the kernels are compiled for 5/3 and mhd_create rejects other values */
  double cfl;              /* effective Courant number; reference 0.5 * 0.9 = 0.45 (diagnostics.hpp:82-84) */
  double cg_tol;           /* relative CG tolerance; reference: Eigen default DBL_EPSILON */
  long long cg_max_iter;   /* <= 0: 2 N, Eigen's default (N = nx * ny corners) */
  int cg_fail_is_error;    /* 0 (reference): a non-converged solve is reported only in the stats */
  int march_rows;          /* rows per work item of the explicit-stage kernel; 0 = auto */
  int bc_x, bc_y;          /* boundary of each axis: MHD_BC_PERIODIC (reference, Grid.hpp:64-80) or
                              MHD_BC_OUTFLOW (zero-gradient; see below) */
} mhd_opts;

/* Boundary conditions.  The reference is periodic only and names the
 * missing piece: a non-periodic run needs another ghost-fill rule
 * (Grid.hpp:141-142).  MHD_BC_OUTFLOW is the zero-gradient rule: along that
 * axis every field is extended beyond the domain by its value at the nearest
 * cell (face) inside -- the ghost fill of the state before each explicit
 * stage, of the recombined field before its resistive term -- and the corner
 * quantities of the resistive term (eta J and the implicit solve's unknown)
 * by the nearest corner, which makes the implicit operator's Laplacian a
 * Neumann one (its Jacobi preconditioner takes the boundary rows' own
 * diagonal).  The CT corner field E_z at the last corner line comes from the
 * ghost-filled fluxes like every other corner.  Supported by the single-run
 * entry points (including mhd_run_resident); batches and the adjoint are
 * periodic. */
#define MHD_BC_PERIODIC 0
#define MHD_BC_OUTFLOW 1

/* The reference's problem setups (include/benchmarks.hpp), for
 * mhd_init_setup / mhd_batch_init_setup. */
#define MHD_SETUP_ORSZAG_TANG 0    /* orszag_tang_initial_conditions, benchmarks.hpp:15-76 */
#define MHD_SETUP_DECAYING_SINE 1  /* decaying_magnetic_sine_wave_initial_conditions, :352-417 */
#define MHD_SETUP_ALFVEN 2         /* alfven_wave_test_initial_values, :78-126 */
#define MHD_SETUP_GAUSSIAN 3       /* gaussian_density_pulse_test_initial_values, :128-176 */
#define MHD_SETUP_ACOUSTIC 4       /* acoustic_wave_test_initial_values, :178-231 */
#define MHD_SETUP_SAUSAGE 5        /* sausage_instability_initial_conditions, :233-293 */
#define MHD_SETUP_KINK 6           /* kink_instability_initial_conditions, :295-350 */

/* Per-step statistics (the reference computes none of these; CG failure is
 * silent there, solvers.hpp:469). */
typedef struct mhd_step_stats {
  int cg_iters[2];         /* Eigen iteration count of the two implicit solves */
  int cg_spmv[2];          /* operator applications actually performed */
  double cg_error[2];      /* sqrt(|r|^2 / |b|^2) at exit (Eigen's tol_error) */
  double dt;               /* the dt the step used */
} mhd_step_stats;

/* Fills *opts with the defaults above. */
MHD_API int mhd_default_opts(mhd_opts* opts);

/* Creates a context for an nx x ny periodic grid with cell sizes dx, dy and
 * resistivity eta (Grid ctor, Grid.hpp:29-54; cell_width/cell_height/
 * resistivity as set by src/integrator_main.cpp:37-39).  opts may be NULL. */
MHD_API int mhd_create(int nx, int ny, double dx, double dy, double eta, const mhd_opts* opts, mhd_ctx** out);
MHD_API int mhd_destroy(mhd_ctx* ctx);
MHD_API const char* mhd_last_error(const mhd_ctx* ctx);

/* Copies the conserved state from padded host arrays (ghosts ignored) and
 * applies the density floor, as the driver's first fillGhostPeriodic +
 * computePrimitives does (src/integrator_main.cpp:52-53, Grid.hpp:115). */
MHD_API int mhd_upload(mhd_ctx* ctx, const double* rho, const double* mx, const double* my,
               const double* E, const double* bxf, const double* byf);

/* Copies the state back into padded host arrays with ghosts filled
 * periodically (Grid::fillGhostPeriodic, Grid.hpp:64-80).  Any pointer may be
 * NULL.  vx, vy, p, bxc, byc are the primitives of Grid::computePrimitives
 * (Grid.hpp:100-136), computed on the device. */
MHD_API int mhd_download(mhd_ctx* ctx, double* rho, double* mx, double* my, double* E,
                 double* bxf, double* byf, double* vx, double* vy, double* p,
                 double* bxc, double* byc);

/* Device-resident counterparts of mhd_upload / mhd_download (no reference
 * counterpart): six compact nx*ny device arrays (HIP device pointers on the
 * context's device, row-major j * nx + i) rho, mx, my, E, bxf, byf.  The
 * upload applies the density floor; NULL entries are skipped on download.
 * Both return after the copies have completed. */
MHD_API int mhd_upload_device(mhd_ctx* ctx, const double* const* fields);
MHD_API int mhd_download_device(mhd_ctx* ctx, double* const* fields);

/* CFL_condition_checker (include/diagnostics.hpp:44-88) on the current state,
 * with the Courant number of opts.cfl. */
MHD_API int mhd_cfl_dt(mhd_ctx* ctx, double* dt);

/* generate_A_matrix + solver.compute (include/solvers.hpp:404-443,
 * src/integrator_main.cpp:58-60): sets alpha of the operator I - alpha L.
 * The driver passes alpha = dt * eta * gamma_ARS. */
MHD_API int mhd_build_operator(mhd_ctx* ctx, double alpha);

/* IMEX (include/solvers.hpp:491-613): one ARS(2,2,2) step of size dt.  The
 * implicit solves use the operator of the last mhd_build_operator call, as
 * the reference's solver does.  stats may be NULL. */
MHD_API int mhd_imex_step(mhd_ctx* ctx, double dt, mhd_step_stats* stats);

/* Unit entry points for parity tests. */
/* explicit_rhs (solvers.hpp:393-402): dU/dt of the current state into padded
 * arrays (physical cells written, ghosts left untouched). */
MHD_API int mhd_explicit_rhs(mhd_ctx* ctx, double* drho, double* dmx, double* dmy, double* dE,
                     double* dbxf, double* dbyf);
/* resistive_B_term (solvers.hpp:445-489) of the current state's face B.
 * implicit != 0 solves with the operator of mhd_build_operator.  iters may be NULL. */
MHD_API int mhd_resistive_term(mhd_ctx* ctx, int implicit, double* bx_rate, double* by_rate, int* iters);
/* The CG solve alone (Eigen ConjugateGradient::solve, solvers.hpp:469) on a
 * compact nx*ny right-hand side. */
MHD_API int mhd_cg_solve(mhd_ctx* ctx, const double* b, double* x, int* iters, double* error);
/* calculate_current_div_B (include/diagnostics.hpp:15-42). */
MHD_API int mhd_divb(mhd_ctx* ctx, double* rms, double* maxabs);

/* Driver loop (src/integrator_main.cpp:52-112): CFL time step (or the
 * constant dt_override when it is non-zero), the clamp of the last step to
 * t_end, the operator rebuild when dt changes by more than 1e-10 relative,
 * then IMEX; stops at t_end or after max_steps steps (max_steps < 0: no
 * limit).  step_dt / cg_iters (2 per step) may be NULL, else hold max_steps
 * entries.  *t_out receives the final time. */
MHD_API int mhd_run(mhd_ctx* ctx, double t_end, double dt_override, int max_steps,
            int* steps_out, double* t_out, double* step_dt, int* cg_iters);

/* Snapshot callback of mhd_run_resident: frame = steps taken (0 before the
 * first step), t = the time after it, dt = that step's dt (frame 0: the
 * initial CFL dt or the override, as the reference's first snapshot records).
 * The stream is idle, so the callback may call mhd_download /
 * mhd_write_snapshot or change the state (the loop continues from the state
 * the callback leaves); it must not call mhd_run_resident on the same
 * context.  A non-zero return stops the loop and is returned. */
typedef int (*mhd_snapshot_fn)(mhd_ctx* ctx, int frame, double t, double dt, void* user);

/* The driver loop of mhd_run (src/integrator_main.cpp:52-112, same
 * arguments and results, identical states, time steps and CG counts) with
 * the per-step decisions on the device: dt, the t_end clamp and the operator
 * rebuild rule are computed by a kernel, and CG solves run a launched
 * iteration budget whose unused passes return at once, so the host reads the
 * driver state once per group of steps instead of three times per step.
 * With snapshot_every > 0 and on_snapshot non-NULL, on_snapshot runs before
 * the first step and after every snapshot_every-th step, like the
 * reference's snapshot cadence (integrator_main.cpp:65-70, 102-111).
 * No reference counterpart beyond the loop itself. */
MHD_API int mhd_run_resident(mhd_ctx* ctx, double t_end, double dt_override, int max_steps, int snapshot_every,
                             mhd_snapshot_fn on_snapshot, void* user, int* steps_out, double* t_out,
                             double* step_dt, int* cg_iters);
/* mhd_run_resident with the reference's snapshot files written while the
 * loop runs on (write_data_to_binary_file, include/diagnostics.hpp:90-131,
 * at the cadence of integrator_main.cpp:65-70, 102-111): before the first
 * step and after every snapshot_every-th step (> 0), the device copies the
 * file's arrays and div B partials into one of two staging buffers, a second
 * stream moves them to pinned host memory, and a writer thread writes
 * path_format with the frame number (one %d conversion, e.g.
 * "output/snapshot_%06d.dat"; header snapshot_rate = snapshot_rate).  The
 * files are byte for byte mhd_write_snapshot's from a mhd_run_resident
 * callback.  Returns after the last file is written: the loop's error, or
 * else the first write error.  Device and pinned host memory: 2 x (7 nx ny
 * + 2 tiles) doubles each, for the call's duration. */
MHD_API int mhd_run_resident_files(mhd_ctx* ctx, double t_end, double dt_override, int max_steps, int snapshot_every,
                                   const char* path_format, int snapshot_rate, int* steps_out, double* t_out,
                                   double* step_dt, int* cg_iters);
/* Tuning of mhd_run_resident: steps enqueued between two read-backs (1..256,
 * default 16), extra CG iterations launched per solve beyond the most of the
 * previous group (default 0), fixed_budget >= 0: that many launched
 * iterations per solve, not adapted (a solve that needs more halts the loop
 * and the host finishes its step; -1 restores the adaptive budget), and
 * graph = 1: pairs of steps are captured once per configuration (B pointers,
 * budgets, dt mode) as a HIP graph and replayed; graph = 2: the even part of
 * each group is captured as one graph (at most 64 steps per graph), one
 * launch per group (0: direct launches). */
MHD_API int mhd_resident_tune(mhd_ctx* ctx, int group_steps, int spec_margin, int fixed_budget, int graph);

/* write_data_to_binary_file (include/diagnostics.hpp:90-131): the current
 * state as a snapshot file in the reference's format -- 9 doubles (nx, ny,
 * ng = 2, snapshot_rate, dx, dy, time_step, rms div B, max |div B|), then the
 * padded (ny+4) x (nx+4) arrays rho, E, p, vx, vy, bxc, byc with ghosts
 * filled (read by scripts/display_functions.py:3-22).  The reference names
 * the file output/snapshot_%06d.dat (the C++ adapter does the same). */
MHD_API int mhd_write_snapshot(mhd_ctx* ctx, const char* path, int snapshot_rate, double time_step);
/* The device generators of mhd_batch_init_fourier / mhd_batch_init_column
 * (declared with the batch entry points) into a context's state. */
MHD_API int mhd_init_fourier(mhd_ctx* ctx, int nmodes, const double* modes);
MHD_API int mhd_init_column(mhd_ctx* ctx, int sausage, int nmodes, const double* amps, const double* phases);
/* Replaces the reference's setup call test_selection(global_state, ...)
 * (src/integrator_main.cpp:50; the switch over the seven setups of
 * include/benchmarks.hpp:419-451, which fill the padded Grid on the host): the setup
 * evaluated per interior cell on the device, at the context's nx, ny, dx, dy
 * (centres (i + 1/2) dx, corners i dx, with the reference's pi =
 * 3.14159265359), straight into its state.  The fields agree with the
 * reference's host loops to a few ulps (the same operations in the same order;
 * the device's sin, cos, exp).  setup: MHD_SETUP_*. */
MHD_API int mhd_init_setup(mhd_ctx* ctx, int setup);
/* Exact-restart files (no reference counterpart: the reference cannot
 * resume): "MHDRST01", 7 doubles (nx, ny, ng, dx, dy, eta, t) and the six
 * padded conserved arrays.  Reading checks the grid and uploads the state;
 * *t (may be NULL) receives the stored time. */
MHD_API int mhd_write_restart(mhd_ctx* ctx, const char* path, double t);
MHD_API int mhd_read_restart(mhd_ctx* ctx, const char* path, double* t);

/* Instrumentation (no reference counterpart).  mask selects kernel ids
 * (bit k = id k, 0 = off, 0x1ff = all); each launch of a selected kernel is
 * bracketed by HIP events on the context's stream, and mhd_profile_get
 * returns the launch count and summed milliseconds per id since the last
 * mhd_profile_enable.  Kernel ids: 0 stage1, 1 stage2, 2 cg_start,
 * 3 cg_pass_a, 4 cg_pass_b, 5 implicit_finalize1, 6 implicit_finalize2,
 * 7 cfl, 8 other. */
MHD_API int mhd_profile_enable(mhd_ctx* ctx, int mask);
MHD_API int mhd_profile_get(mhd_ctx* ctx, int kernel_id, long long* launches, double* total_ms);
/* Bytes-calibration helper for hardware counters: one plain device copy
 * kernel of nbytes (rounded down to doubles), run synchronously. */
MHD_API int mhd_calibration_copy(mhd_ctx* ctx, long long nbytes);
/* Streaming reference for a read/write mix (no reference counterpart): a
 * plain kernel that reads n_in and writes n_out arrays of n doubles once per
 * element, launched twice untimed and `reps` times between HIP events on the
 * context's stream; *ms_per_launch receives the mean.  Mixes (n_in, n_out):
 * (1,1) (2,1) (2,2) (3,2) (7,2) (6,12) (12,6) -- those of the step's kernels.
 * Allocates and frees its (n_in + n_out) n doubles. */
MHD_API int mhd_calibration_mix(mhd_ctx* ctx, int n_in, int n_out, long long n, int reps, double* ms_per_launch);
MHD_API int mhd_synchronize(mhd_ctx* ctx);

/* Allocation-free CFL time step.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops that forbid hipMalloc).  The time step of mhd_cfl_dt
 * (CFL_condition_checker, diagnostics.hpp:44-88; Courant number cfl, the
 * reference's 0.45 reproduced bit for bit) of a state the caller holds in
 * six compact nx*ny device arrays (rho, mx, my, E, bxf, byf, row-major
 * j*nx+i, the context layout of mhd_upload_device), written to the device
 * double *dt_device (NaN if a wave speed is not finite).  No context, no
 * allocation, no host synchronisation: two kernels on `stream` (a
 * hipStream_t; NULL: the null stream).  Scratch: a 64 KiB file-scope device
 * array per GPU, held until the process ends, shared by every call on that
 * GPU -- calls that overlap on one GPU race.  Limit: the per-tile partials
 * of the grid must fit the scratch (at most 4096 tiles of 256 columns:
 * nx <= 1048576), nx, ny >= 4, nx * ny <= 2^30; larger grids return
 * MHD_EINVAL (mhd_last_error(NULL) says why) without launching. */
MHD_API int mhd_cfl_dt_noalloc(int nx, int ny, double dx, double dy, double eta, double cfl, int bc_x, int bc_y,
                               const double* const* fields, double* dt_device, void* stream);

/* ------------------------------------------------------------------------
 * Discrete adjoint (no reference counterpart: the reference has no
 * derivatives).  The adjoint of the whole IMEX step: reverse-mode
 * primitives / MUSCL-minmod / HLL / upwind CT at the recorded forward state
 * (the minmod, HLL and CT branch choices of the forward pass), transposed
 * CT and resistive stencils, density and pressure floors as 0/1 masks, and
 * for each implicit solve a second CG solve with the same symmetric operator
 * (I - alpha Lap).  Requires eta > 0 and a constant time step.
 * ------------------------------------------------------------------------ */

/* Runs nsteps IMEX steps of constant dt from the current state (operator
 * alpha = dt eta gamma, as the driver builds it for a fixed dt,
 * src/integrator_main.cpp:56-60), keeping a copy of the state every
 * checkpoint_every steps.  The context then holds the final state. */
MHD_API int mhd_adjoint_forward(mhd_ctx* ctx, double dt, int nsteps, int checkpoint_every);
/* Given lam_final = dJ/dU(final) (six padded arrays rho, mx, my, E, bxf, byf),
 * returns lam0 = dJ/dU(initial) (six padded arrays, ghosts filled; the array
 * pointer or any entry may be NULL) and *eta_bar = dJ/d eta through every
 * step of the last mhd_adjoint_forward.  Each checkpoint segment is
 * recomputed once; the context's state is left at the final state. */
MHD_API int mhd_adjoint_backward(mhd_ctx* ctx, const double* const* lam_final, double* const* lam0, double* eta_bar);
/* mhd_adjoint_backward on compact nx*ny device arrays (HIP device pointers on
 * the context's device) instead of padded host arrays; for callers that keep
 * fields on the GPU (the LibTorch autograd Function of this package). */
MHD_API int mhd_adjoint_backward_device(mhd_ctx* ctx, const double* const* lam_final, double* const* lam0,
                                double* eta_bar);
/* Launch mode of the backward pass (no reference counterpart).  on != 0
 * (default): a pass with no host round trip (Jacobi adjoint solves) is
 * captured once per configuration as a HIP graph and replayed; on == 0:
 * its kernels are launched directly, in the same order with the same
 * arguments (bit-identical results).  Turn it off under tools that trace
 * graph-launched kernels unreliably. */
MHD_API int mhd_adjoint_use_graph(mhd_ctx* ctx, int on);
/* Branch-frozen forward map, for checking the adjoint against finite
 * differences (no reference counterpart).  The step is piecewise smooth: the
 * minmod choice (MUSCL.hpp:11-17), the HLL case and its wave-speed min / max
 * (HLL.hpp:145-152), the upwind corner E_z (solvers.hpp:326-338) and the
 * density / pressure floors (Grid.hpp:115,124) switch between smooth
 * branches, and the adjoint differentiates each with the branch taken at the
 * recorded forward state.  With on != 0 the next mhd_adjoint_backward also
 * records those choices (one 64-bit word per cell and stage, 2 nsteps nx ny
 * words on the device; the pass then runs without graph replay and with the
 * unfused right-hand-side adjoint). */
MHD_API int mhd_adjoint_record_branches(mhd_ctx* ctx, int on);
/* Runs nsteps (<= the recorded steps) IMEX steps of the recorded constant dt
 * from the current state with every branch choice read from the record, the
 * implicit solves exact to rounding (Jacobi sweeps to below 2^-53 of the
 * solution; operators with 2 (cx + cy) / (1 + 2 (cx + cy)) > 1/4 are refused)
 * and resistivity eta (<= 0: the context's).  A smooth function of the
 * initial state and eta whose derivative is what the adjoint computes; equal
 * to the forward step wherever the state makes the recorded choices.  The
 * context then holds the final state. */
MHD_API int mhd_adjoint_forward_frozen(mhd_ctx* ctx, int nsteps, double eta);
/* Unit entry point: out = L'(U)^T lam for the explicit right-hand side L of
 * explicit_rhs (solvers.hpp:393-402) at the current state U (padded arrays;
 * out's physical cells are written), by the fused single-pass kernel (the
 * backward pass uses it from 2^19 cells on and the unfused kernels below). */
MHD_API int mhd_explicit_rhs_vjp(mhd_ctx* ctx, const double* const* lam, double* const* out);
/* The same product by four unfused gather kernels (primitives, face
 * adjoints, MUSCL / primitive adjoint, face-field terms); the reference the
 * fused kernel of mhd_explicit_rhs_vjp is tested against. */
MHD_API int mhd_explicit_rhs_vjp_unfused(mhd_ctx* ctx, const double* const* lam, double* const* out);

/* ------------------------------------------------------------------------
 * Batched ensembles (no reference counterpart: the reference advances one
 * Grid per process).  R independent runs, each with its own size,
 * cell size and resistivity, advance together: one launch per kernel covers
 * every run, and each run follows the reference driver loop
 * (src/integrator_main.cpp:76-112) on the device -- its own CFL time step
 * (opts.cfl), the clamp of its last step to its t_end, the operator rebuild
 * when its dt changes by more than 1e-10 relative -- and the IMEX step of
 * mhd_imex_step.  Each run's arithmetic is that of a single-run context.
 * ------------------------------------------------------------------------ */
typedef struct mhd_batch mhd_batch;

/* nx, ny, dx, dy, eta: arrays of nruns entries.  Every run starts at t = 0
 * with no end time and no step limit (see mhd_batch_set_schedule).  The
 * boundaries of opts (bc_x, bc_y: periodic or MHD_BC_OUTFLOW) apply to every
 * run; the batched adjoint (mhd_batch_adjoint_*) takes periodic runs only. */
MHD_API int mhd_batch_create(int nruns, const int* nx, const int* ny, const double* dx, const double* dy,
                     const double* eta, const mhd_opts* opts, mhd_batch** out);
MHD_API int mhd_batch_destroy(mhd_batch* b);
MHD_API const char* mhd_batch_last_error(const mhd_batch* b);
/* Per-run end time, step limit (< 0: none) and constant time step (0: CFL);
 * any array may be NULL to keep the current values. */
MHD_API int mhd_batch_set_schedule(mhd_batch* b, const double* t_end, const int* max_steps, const double* dt_fixed);
/* State of one run from / to padded host arrays (as mhd_upload / mhd_download);
 * an upload restarts that run's clock and step count. */
MHD_API int mhd_batch_upload(mhd_batch* b, int run, const double* rho, const double* mx, const double* my,
                     const double* E, const double* bxf, const double* byf);
MHD_API int mhd_batch_download(mhd_batch* b, int run, double* rho, double* mx, double* my, double* E,
                       double* bxf, double* byf);
/* Initial states of the ensemble workloads generated on the device, straight
 * into a run (no host arrays, no upload; the run restarts as after an upload).
 * No reference counterpart: the reference's setups are host loops
 * (include/benchmarks.hpp:15-417); these evaluate the seeded setups of the
 * ensemble workloads (the Python package's setups.fourier_ensemble_member and
 * setups.kink_column / sausage_column, which keep the seeded draws) per cell.
 * Square runs only (nx == ny, dx == dy).
 *   fourier: rho = p = 1, B = (1, 0.5) + curl of a corner potential, v and
 *     the potential from nmodes modes of 8 doubles each: kx, ky (nonzero
 *     wave vector), phases of vx, vy, Az, amplitudes of vx, vy, Az; on a
 *     box of side nx * dx.
 *   column: the kink (sausage = 0) or sausage (sausage = 1) column on the
 *     unit box (dx = 1 / nx) with displacement sum_k amps[k]
 *     sin(2 pi (k + 1) y + phases[k]).
 * The fields agree with the host generators to a few ulps (the same
 * operations in the same order; the device's sin, cos, exp). */
MHD_API int mhd_batch_init_fourier(mhd_batch* b, int run, int nmodes, const double* modes);
MHD_API int mhd_batch_init_column(mhd_batch* b, int run, int sausage, int nmodes, const double* amps,
                                  const double* phases);
/* mhd_init_setup into run `run` of a batch (at the run's nx, ny, dx, dy; the
 * run restarts as after an upload). */
MHD_API int mhd_batch_init_setup(mhd_batch* b, int run, int setup);
/* One driver step of every run that has not reached its end time or step
 * limit; *n_active (may be NULL) receives the number of runs that stepped. */
MHD_API int mhd_batch_step(mhd_batch* b, int* n_active);
/* Steps every run until none is active or max_batch_steps batch steps (< 0:
 * no limit) have been taken; *steps_done: the batch steps in which some run
 * stepped.  In group mode (mhd_batch_tune, the default: groups of 16) the
 * steps are enqueued in groups with the per-run driver on the device and one
 * read-back of the driver state per group: each CG solve runs a launched
 * iteration budget (the longest solve of that slot in the last group plus a
 * margin; a solve's passes after its convergence return at once), a check
 * after each solve halts the group if a run's solve is unfinished, and the
 * host then finishes that solve with host-checked iterations and resumes.
 * The same kernels in the same order as mhd_batch_step: states, time steps
 * and iteration counts are those of calling mhd_batch_step repeatedly.  With
 * group_steps = 0, mhd_batch_step in a loop. */
MHD_API int mhd_batch_run(mhd_batch* b, int max_batch_steps, int* steps_done);
/* Launch mode of mhd_batch_run: group_steps (0: one host-checked step at a
 * time; 1..4096: steps per group), spec_margin (extra launched CG iterations
 * per solve), graph (0: direct launches; 1: pairs of steps captured once as a
 * hipGraph and replayed; 2: the even part of each group captured as one
 * hipGraph (at most 64 steps per graph) and launched once per group; graphs
 * are kept per step count,
 * B-buffer parity, budgets and fusion), fused_cfl (1: the final
 * recombination also forms the next step's CFL partials; 0: a separate CFL
 * pass at the start of each step).  Results are identical in every mode. */
MHD_API int mhd_batch_tune(mhd_batch* b, int group_steps, int spec_margin, int graph, int fused_cfl);
/* The batch's current launch shape and mode (any pointer may be NULL). */
MHD_API int mhd_batch_get_tuning(mhd_batch* b, int* stage_rows, int* tile_rows, int* group_steps, int* graph,
                                 int* fused_cfl);
/* Autotuner: times trial_steps steps of the uploaded batch (in the configured
 * launch mode) for each candidate launch shape (rows per explicit-stage
 * segment x rows per CG / finalize tile, each in {16, 32, 64}), then, in
 * group mode with a resistive run, each stage fusion (CFL partials fused into
 * the final recombination or a separate pass) x launch mode (direct launches,
 * captured pairs or whole captured groups) with the best shape; restores the
 * state after every trial
 * and keeps the fastest configuration.  The chosen rows are returned (either
 * pointer may be NULL; mhd_batch_get_tuning gives the rest).  Results do not
 * depend on the configuration beyond the summation order of the CG dot
 * products (the tile shape). */
MHD_API int mhd_batch_autotune(mhd_batch* b, int trial_steps, int* stage_rows, int* tile_rows);
/* Steps taken, time, last dt and the two CG iteration counts of the last step of one run. */
MHD_API int mhd_batch_run_info(mhd_batch* b, int run, int* steps, double* t, double* last_dt, int* cg_iters);
MHD_API int mhd_batch_synchronize(mhd_batch* b);

/* ------------------------------------------------------------------------
 * Batched adjoint (no reference counterpart): mhd_adjoint_* for every run of
 * a batch at once.  Every run needs eta > 0 and a fixed time step
 * (mhd_batch_set_schedule dt_fixed > 0); each run's gradient is that of a
 * single-run context on the same inputs (Jacobi or CG adjoint solves by the
 * same rule; sums of the eta derivative in another order).
 * ------------------------------------------------------------------------ */

/* Runs nsteps steps of every run from the uploaded states, recording each
 * step's stage data and keeping a copy of the state every checkpoint_every
 * steps (the stage data of one segment of checkpoint_every steps is held on
 * the device; checkpoint_every >= nsteps keeps all of it and the backward pass
 * recomputes nothing).  Fails if a run reaches its t_end or step limit on the
 * way.  The batch then holds the final states.  Sets every run's final-state
 * adjoint to zero. */
MHD_API int mhd_batch_adjoint_forward(mhd_batch* b, int nsteps, int checkpoint_every);
/* dJ/dU(final) of one run: six padded arrays rho, mx, my, E, bxf, byf. */
MHD_API int mhd_batch_adjoint_set_lambda(mhd_batch* b, int run, const double* const* lam_final);
/* The backward pass of every run through the steps of the last
 * mhd_batch_adjoint_forward.  The final-state adjoints are consumed: set them
 * again before another backward pass. */
MHD_API int mhd_batch_adjoint_backward(mhd_batch* b);
/* dJ/dU(initial) of one run (six padded arrays, ghosts filled; the array
 * pointer or any entry may be NULL) and dJ/d eta (may be NULL). */
MHD_API int mhd_batch_adjoint_result(mhd_batch* b, int run, double* const* lam0, double* eta_bar);

#ifdef __cplusplus
}
#endif
#endif
