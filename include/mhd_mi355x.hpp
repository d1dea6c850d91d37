// mhd_mi355x.hpp -- C++ drop-in adapter over the C ABI (mhd_mi355x.h).
//
// Host code written against the reference's step interface
// (RikibabManos/Magnetohydrodynamics_Simulator, include/solvers.hpp and
// include/diagnostics.hpp) keeps its Grid and its call sites; only the
// scratch/solver arguments change.  GridT is any type with the reference
// Grid's members (Grid.hpp:9-26): nx, ny, cell_width, cell_height,
// resistivity and std::vector<double>-like rho, mx, my, E, bxf, byf, vx, vy,
// p, bxc, byc in the padded (nx + 4) x (ny + 4) row-major layout.
//
//   reference                                          this adapter
//   IMEX(g, g1, g2, mem, RV, tableau, solver, dt)       mhd_mi355x::IMEX(g, ctx, dt)
//   explicit_rhs(g, dU, RV, mem)                        mhd_mi355x::explicit_rhs(g, ctx, d...)
//   resistive_B_term(g, bx, by, solver, mem, impl)      mhd_mi355x::resistive_B_term(g, ctx, bx, by, impl)
//   generate_A_matrix(g, mem, alpha); solver.compute    mhd_mi355x::generate_A_matrix(ctx, alpha)
//   CFL_condition_checker(g)                            mhd_mi355x::CFL_condition_checker(g, ctx)
//   calculate_current_div_B(g)                          mhd_mi355x::calculate_current_div_B(g, ctx)
//   write_data_to_binary_file(g, frame, rate, dt)       mhd_mi355x::write_data_to_binary_file(g, ctx, frame, rate, dt)
//   test_selection(g, mem) (benchmarks.hpp:419-451)     mhd_mi355x::test_selection(g, ctx, choice)
//   main's time loop (integrator_main.cpp:52-112)      mhd_mi355x::integrate(g, ctx, t_end, dt_override, rate)
//
// The Grid-taking functions copy the state to the device, run, and copy it
// back (PCIe both ways).  Loops that keep the state on the GPU call
// Context::upload once, Context::step / Context::cfl_dt per step and
// Context::download when they need host data.
#pragma once

#include <cstdio>
#include <stdexcept>
#include <string>

#include "mhd_mi355x.h"

namespace mhd_mi355x {

class Error : public std::runtime_error {
 public:
  Error(int code, const std::string& m) : std::runtime_error(m), code(code) {}
  int code;
};

inline void check(int rc, const mhd_ctx* c) {
  if (rc != MHD_OK) throw Error(rc, mhd_last_error(c));
}

struct divergence_checker_values {   // diagnostics.hpp:8-13
  double root_mean_square_div;
  double max_norm_divergence;
};

class Context {
 public:
  Context(int nx, int ny, double dx, double dy, double eta, const mhd_opts* o = nullptr) {
    check(mhd_create(nx, ny, dx, dy, eta, o, &c_), nullptr);
  }
  template <class GridT>
  explicit Context(const GridT& g, const mhd_opts* o = nullptr)
      : Context(g.nx, g.ny, g.cell_width, g.cell_height, g.resistivity, o) {}
  ~Context() {
    if (c_) mhd_destroy(c_);
  }
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;

  mhd_ctx* get() const { return c_; }

  template <class GridT>
  void upload(const GridT& g) {
    check(mhd_upload(c_, g.rho.data(), g.mx.data(), g.my.data(), g.E.data(), g.bxf.data(), g.byf.data()), c_);
  }
  // Conserved fields, primitives and cell-centred B, ghosts filled (the
  // state fillGhostPeriodic + computePrimitives leave behind).
  template <class GridT>
  void download(GridT& g) {
    check(mhd_download(c_, g.rho.data(), g.mx.data(), g.my.data(), g.E.data(), g.bxf.data(), g.byf.data(),
                       g.vx.data(), g.vy.data(), g.p.data(), g.bxc.data(), g.byc.data()),
          c_);
  }
  double cfl_dt() {
    double dt = 0;
    check(mhd_cfl_dt(c_, &dt), c_);
    return dt;
  }
  mhd_step_stats step(double dt) {
    mhd_step_stats s{};
    check(mhd_imex_step(c_, dt, &s), c_);
    return s;
  }

 private:
  mhd_ctx* c_ = nullptr;
};

inline void generate_A_matrix(Context& ctx, double alpha) { check(mhd_build_operator(ctx.get(), alpha), ctx.get()); }

template <class GridT>
void IMEX(GridT& grid_imex, Context& ctx, double dt, mhd_step_stats* stats = nullptr) {
  ctx.upload(grid_imex);
  mhd_step_stats s = ctx.step(dt);
  ctx.download(grid_imex);
  if (stats) *stats = s;
}

template <class GridT, class Vec>
void explicit_rhs(GridT& g, Context& ctx, Vec& rho_dot, Vec& mx_dot, Vec& my_dot, Vec& E_dot, Vec& bx_dot,
                  Vec& by_dot) {
  ctx.upload(g);
  check(mhd_explicit_rhs(ctx.get(), rho_dot.data(), mx_dot.data(), my_dot.data(), E_dot.data(), bx_dot.data(),
                         by_dot.data()),
        ctx.get());
}

template <class GridT, class Vec>
void resistive_B_term(const GridT& g, Context& ctx, Vec& bx_rate, Vec& by_rate, bool solve_implicitly) {
  ctx.upload(g);
  check(mhd_resistive_term(ctx.get(), solve_implicitly ? 1 : 0, bx_rate.data(), by_rate.data(), nullptr), ctx.get());
}

// test_selection (benchmarks.hpp:419-451) with the menu choice as an argument
// in place of the prompt on std::cin: 1 Orszag-Tang, 2 sausage, 3 kink,
// 4 decaying sine, 5 Alfven wave, 6 Gaussian pulse, 7 acoustic wave; any
// other value runs Orszag-Tang, as the reference's default does.  The setup
// is generated on the device into ctx (the context's grid: g's nx, ny,
// cell_width, cell_height) and downloaded into g.
inline int setup_of_choice(int choice) {
  switch (choice) {
    case 2: return MHD_SETUP_SAUSAGE;
    case 3: return MHD_SETUP_KINK;
    case 4: return MHD_SETUP_DECAYING_SINE;
    case 5: return MHD_SETUP_ALFVEN;
    case 6: return MHD_SETUP_GAUSSIAN;
    case 7: return MHD_SETUP_ACOUSTIC;
    default: return MHD_SETUP_ORSZAG_TANG;
  }
}

template <class GridT>
void test_selection(GridT& g, Context& ctx, int choice) {
  check(mhd_init_setup(ctx.get(), setup_of_choice(choice)), ctx.get());
  ctx.download(g);
}

template <class GridT>
double CFL_condition_checker(const GridT& g, Context& ctx) {
  ctx.upload(g);
  return ctx.cfl_dt();
}

template <class GridT>
divergence_checker_values calculate_current_div_B(const GridT& g, Context& ctx) {
  ctx.upload(g);
  divergence_checker_values v{};
  check(mhd_divb(ctx.get(), &v.root_mean_square_div, &v.max_norm_divergence), ctx.get());
  return v;
}

// write_data_to_binary_file (diagnostics.hpp:90-131): the reference's file
// name output/snapshot_<frame, 6 digits>.dat and format.
template <class GridT>
void write_data_to_binary_file(const GridT& g, Context& ctx, int current_frame_index, int snapshot_rate,
                               double time_step) {
  ctx.upload(g);
  char name[64];
  std::snprintf(name, sizeof(name), "output/snapshot_%06d.dat", current_frame_index);
  check(mhd_write_snapshot(ctx.get(), name, snapshot_rate, time_step), ctx.get());
}

// The time loop of the reference's main (src/integrator_main.cpp:52-112) on
// the device: g's state is uploaded once, advanced to t_end (constant dt when
// dt_override != 0) by mhd_run_resident, with a snapshot
// output/snapshot_<frame>.dat before the first step and every
// snapshot_frequency steps (0: none), written while the loop runs on
// (mhd_run_resident_files), and downloaded into g at the end.  Returns the
// number of steps taken.
template <class GridT>
int integrate(GridT& g, Context& ctx, double t_end, double dt_override, int snapshot_frequency) {
  ctx.upload(g);
  int steps = 0;
  double t = 0.0;
  if (snapshot_frequency > 0)
    check(mhd_run_resident_files(ctx.get(), t_end, dt_override, -1, snapshot_frequency, "output/snapshot_%06d.dat",
                                 snapshot_frequency, &steps, &t, nullptr, nullptr),
          ctx.get());
  else
    check(mhd_run_resident(ctx.get(), t_end, dt_override, -1, 0, nullptr, nullptr, &steps, &t, nullptr, nullptr),
          ctx.get());
  ctx.download(g);
  return steps;
}

}  // namespace mhd_mi355x
