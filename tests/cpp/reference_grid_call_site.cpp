// A call site that drives the GPU path the way the reference's main drives
// its own step (src/integrator_main.cpp:34-39, 52-112), on the reference's
// own Grid (include/Grid.hpp, taken from the reference tree at build time)
// through this library's C++ adapter (include/mhd_mi355x.hpp).
//
//   reference_grid_call_site IN OUT [SNAPSHOT_FREQUENCY]
//   reference_grid_call_site --setup CHOICE NX NY ETA T_END OUT [SNAPSHOT_FREQUENCY]
//
// IN: doubles nx, ny, dx, dy, eta, t_end, then the six padded (ny+4) x (nx+4)
// arrays rho, mx, my, E, bxf, byf.  With --setup the Grid is set up as main
// does (integrator_main.cpp:34-39: cell sizes 2 pi / n with the reference's
// pi) and filled by test_selection with the menu choice CHOICE (1-7,
// benchmarks.hpp:419-451).  The program then calls, in this order:
// explicit_rhs of the initial state,
// CFL_condition_checker, generate_A_matrix for that time step, one IMEX step,
// resistive_B_term (implicit) and calculate_current_div_B of the stepped
// state, then integrate to t_end (snapshot files output/snapshot_<frame>.dat
// every SNAPSHOT_FREQUENCY steps; none by default).
// OUT: doubles dt, steps of the integration, rms and max |div B| after the
// step, then padded arrays: the six explicit right-hand sides, the two
// resistive rates, the six fields after the IMEX step, the six at t_end, the
// six of the initial state.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "Grid.hpp"
#include "mhd_mi355x.hpp"

namespace {

bool read_doubles(std::FILE* f, double* p, size_t n) { return std::fread(p, sizeof(double), n, f) == n; }

std::vector<double>* field(Grid& g, int q) {
  std::vector<double>* f[6] = {&g.rho, &g.mx, &g.my, &g.E, &g.bxf, &g.byf};
  return f[q];
}

}  // namespace

int main(int argc, char** argv) {
  const bool from_setup = argc >= 2 && std::string(argv[1]) == "--setup";
  const int base = from_setup ? 7 : 2;   // argv index of OUT
  if (argc != base + 1 && argc != base + 2) {
    std::fprintf(stderr, "usage: %s IN OUT [SNAPSHOT_FREQUENCY]\n       %s --setup CHOICE NX NY ETA T_END OUT "
                 "[SNAPSHOT_FREQUENCY]\n", argv[0], argv[0]);
    return 2;
  }
  const int snapshot_frequency = argc == base + 2 ? std::atoi(argv[base + 1]) : 0;
  int nx, ny, choice = 0;
  double dx, dy, eta, t_end;
  std::FILE* in = nullptr;
  if (from_setup) {
    choice = std::atoi(argv[2]);
    nx = std::atoi(argv[3]);
    ny = std::atoi(argv[4]);
    eta = std::atof(argv[5]);
    t_end = std::atof(argv[6]);
    constexpr double pi = 3.14159265359;   // integrator_main.cpp:36-38
    dx = 2 * pi / nx;
    dy = 2 * pi / ny;
  } else {
    in = std::fopen(argv[1], "rb");
    if (!in) return 2;
    double hdr[6];
    if (!read_doubles(in, hdr, 6)) return 2;
    nx = (int)hdr[0];
    ny = (int)hdr[1];
    dx = hdr[2];
    dy = hdr[3];
    eta = hdr[4];
    t_end = hdr[5];
  }
  Grid grid(nx, ny, 2);
  grid.cell_width = dx;
  grid.cell_height = dy;
  grid.resistivity = eta;
  if (in) {
    for (int q = 0; q < 6; ++q)
      if (!read_doubles(in, field(grid, q)->data(), field(grid, q)->size())) return 2;
    std::fclose(in);
  }

  std::FILE* out = std::fopen(argv[base], "wb");
  if (!out) return 2;
  try {
    mhd_opts o;
    mhd_default_opts(&o);
    o.cfl = 0.4;
    o.cg_tol = 1e-10;
    mhd_mi355x::Context ctx(grid, &o);
    if (from_setup) mhd_mi355x::test_selection(grid, ctx, choice);   // integrator_main.cpp:50
    std::vector<double> initial[6];
    for (int q = 0; q < 6; ++q) initial[q] = *field(grid, q);
    std::vector<double> rhs[8];
    for (auto& v : rhs) v.assign(grid.rho.size(), 0.0);
    mhd_mi355x::explicit_rhs(grid, ctx, rhs[0], rhs[1], rhs[2], rhs[3], rhs[4], rhs[5]);
    const double dt = mhd_mi355x::CFL_condition_checker(grid, ctx);
    const double gamma_ars = (2.0 - std::sqrt(2.0)) / 2.0;   // IMEX_Butcher_Tableus::gamma (solvers.hpp:41)
    mhd_mi355x::generate_A_matrix(ctx, (dt * grid.resistivity) * gamma_ars);   // integrator_main.cpp:58-59
    mhd_mi355x::IMEX(grid, ctx, dt);
    std::vector<double> after_step[6];
    for (int q = 0; q < 6; ++q) after_step[q] = *field(grid, q);
    mhd_mi355x::resistive_B_term(grid, ctx, rhs[6], rhs[7], true);
    const mhd_mi355x::divergence_checker_values div = mhd_mi355x::calculate_current_div_B(grid, ctx);
    const int steps = mhd_mi355x::integrate(grid, ctx, t_end, 0.0, snapshot_frequency);
    const double head[4] = {dt, (double)steps, div.root_mean_square_div, div.max_norm_divergence};
    std::fwrite(head, sizeof(double), 4, out);
    for (int q = 0; q < 8; ++q) std::fwrite(rhs[q].data(), sizeof(double), rhs[q].size(), out);
    for (int q = 0; q < 6; ++q) std::fwrite(after_step[q].data(), sizeof(double), after_step[q].size(), out);
    for (int q = 0; q < 6; ++q) std::fwrite(field(grid, q)->data(), sizeof(double), field(grid, q)->size(), out);
    for (int q = 0; q < 6; ++q) std::fwrite(initial[q].data(), sizeof(double), initial[q].size(), out);
  } catch (const mhd_mi355x::Error& e) {
    std::fprintf(stderr, "mhd error %d: %s\n", e.code, e.what());
    std::fclose(out);
    return 1;
  }
  std::fclose(out);
  return 0;
}
