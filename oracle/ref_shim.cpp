// ref_shim.cpp -- C entry points over the reference's own Eigen-free headers
// (TEST INFRASTRUCTURE; built into oracle/_ref/ by oracle/Makefile).
//
// The reference's include/Grid.hpp, include/MUSCL.hpp, include/HLL.hpp and
// include/diagnostics.hpp compile on their own; include/solvers.hpp and
// include/benchmarks.hpp need Eigen, which this machine does not have, so they
// are not built.  This file holds no reference code: it includes the headers
// from the reference tree (-I .../include) and exposes a few of their functions
// with plain pointers so tests can compare the oracle against them bit for bit.
#include <cstring>
#include <vector>

#include "Grid.hpp"
#include "MUSCL.hpp"
#include "HLL.hpp"
#include "diagnostics.hpp"

namespace {

// Field order: rho, mx, my, E, bxf, byf, vx, vy, p, bxc, byc, Ez.
void load(Grid& g, double* const* f) {
  std::vector<double>* dst[12] = {&g.rho, &g.mx, &g.my, &g.E, &g.bxf, &g.byf,
                                  &g.vx,  &g.vy, &g.p,  &g.bxc, &g.byc, &g.Ez};
  for (int k = 0; k < 12; ++k)
    if (f[k]) std::memcpy(dst[k]->data(), f[k], dst[k]->size() * sizeof(double));
}

void store(const Grid& g, double* const* f) {
  const std::vector<double>* src[12] = {&g.rho, &g.mx, &g.my, &g.E, &g.bxf, &g.byf,
                                        &g.vx,  &g.vy, &g.p,  &g.bxc, &g.byc, &g.Ez};
  for (int k = 0; k < 12; ++k)
    if (f[k]) std::memcpy(f[k], src[k]->data(), src[k]->size() * sizeof(double));
}

Grid make(int nx, int ny, double dx, double dy, double eta, double* const* f) {
  Grid g(nx, ny, 2);
  g.cell_width = dx;
  g.cell_height = dy;
  g.resistivity = eta;
  load(g, f);
  return g;
}

void to_prim(const double* a, PrimitiveState& s) {
  s.rho = a[0]; s.vx = a[1]; s.vy = a[2]; s.p = a[3]; s.bx = a[4]; s.by = a[5];
}

void from_flux(const FluxVector& v, double* o) {
  o[0] = v.rho; o[1] = v.mx; o[2] = v.my; o[3] = v.E; o[4] = v.bx; o[5] = v.by;
}

}  // namespace

extern "C" {

// Grid::fillGhostPeriodic() then Grid::computePrimitives(), in place on 12 padded arrays.
void ref_fill_and_primitives(int nx, int ny, double* const* f) {
  Grid g = make(nx, ny, 1.0, 1.0, 0.0, f);
  g.fillGhostPeriodic();
  g.computePrimitives();
  store(g, f);
}

// MUSCLReconstructAll() into 20 padded arrays in ReconstructedValues member order.
void ref_muscl_all(int nx, int ny, double* const* f, double* const* rv) {
  Grid g = make(nx, ny, 1.0, 1.0, 0.0, f);
  ReconstructedValues R(static_cast<size_t>(g.nxt) * g.nyt);
  MUSCLReconstructAll(g, R);
  std::vector<double>* out[20] = {
      &R.rhoXFaceL, &R.rhoXFaceR, &R.rhoYFaceL, &R.rhoYFaceR, &R.vxXFaceL, &R.vxXFaceR, &R.vxYFaceL,
      &R.vxYFaceR,  &R.vyXFaceL,  &R.vyXFaceR,  &R.vyYFaceL,  &R.vyYFaceR, &R.pXFaceL,  &R.pXFaceR,
      &R.pYFaceL,   &R.pYFaceR,   &R.bxYFaceL,  &R.bxYFaceR,  &R.byXFaceL, &R.byXFaceR};
  for (int k = 0; k < 20; ++k) std::memcpy(rv[k], out[k]->data(), out[k]->size() * sizeof(double));
}

double ref_minmod(double a, double b) { return minmod(a, b); }

double ref_fast_speed(const double* prim, double bn) {
  PrimitiveState s;
  to_prim(prim, s);
  return getFastWaveSpeed(s, bn);
}

void ref_hll_x(const double* L, const double* R, double* out) {
  PrimitiveState l, r;
  to_prim(L, l);
  to_prim(R, r);
  from_flux(computeHLLFluxX(l, r), out);
}

void ref_hll_y(const double* D, const double* U, double* out) {
  PrimitiveState d, u;
  to_prim(D, d);
  to_prim(U, u);
  from_flux(computeHLLFluxY(d, u), out);
}

// CFL_condition_checker() on a grid whose primitives are already computed.
double ref_cfl(int nx, int ny, double dx, double dy, double eta, double* const* f) {
  Grid g = make(nx, ny, dx, dy, eta, f);
  return CFL_condition_checker(g);
}

void ref_divb(int nx, int ny, double dx, double dy, double* const* f, double* rms, double* mx) {
  Grid g = make(nx, ny, dx, dy, 0.0, f);
  divergence_checker_values v = calculate_current_div_B(g);
  *rms = v.root_mean_square_div;
  *mx = v.max_norm_divergence;
}

}  // extern "C"
