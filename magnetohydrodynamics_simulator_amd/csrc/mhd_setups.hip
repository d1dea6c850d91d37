// mhd_setups.hip -- initial states generated on the device: the reference's
// problem setups and those of the ensemble workloads (C ABI
// mhd_init_setup, mhd_batch_init_setup, mhd_batch_init_fourier,
// mhd_batch_init_column, mhd_init_fourier, mhd_init_column).
//
// The reference's setups are host loops over a padded Grid
// (include/benchmarks.hpp:15-417); the seeded setups of the ensemble
// workloads are host functions of the Python package as well
// (setups.fourier_ensemble_member, setups.kink_column / sausage_column).
// For ensembles of hundreds of runs the host evaluation and the upload of
// every field dominate the set-up, so these kernels evaluate the same
// formulas per cell straight into a batch run's pooled arrays (or a
// context's fields).  The host keeps the seeded draws -- the Fourier mode
// table, the column amplitudes and phases -- which define the ensemble.
//
// The arithmetic follows the host generators operation by operation (same
// coordinates, same order of sums and products, no contraction into FMAs);
// the transcendental functions are the device's, so the fields agree with the
// host's to a few ulps rather than bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/mhd_mi355x.h"
#include "mhd_batch.hpp"
#include "mhd_context.hpp"
#include "mhd_kernels.hpp"

namespace mhd {

namespace {

constexpr double kTwoPi = 6.283185307179586;        // 2 * numpy.pi, as the host generators form it
constexpr double kGammaM1 = 5.0 / 3.0 - 1.0;        // GAMMA - 1.0 of ConservedState::fromPrimitive

// One Fourier mode of an ensemble member: wave numbers, phases and
// amplitudes of vx, vy and the vector potential.
struct FourierMode {
  double kx, ky, ph_vx, ph_vy, ph_az, a_vx, a_vy, a_az;
};

struct InitArgs {
  double* f[6];   // rho, mx, my, E, bxf, byf of the run (compact, periodic)
  int nx, ny;
  double dx, dy;
};

// ConservedState::fromPrimitive (HLL.hpp:31-47) with face B, as the setups use it.
__device__ __forceinline__ void store_state(const InitArgs& a, long long o, double rho, double vx, double vy,
                                            double p, double bx, double by) {
#pragma clang fp contract(off)
  const double v_sq = vx * vx + vy * vy;
  const double b_sq = bx * bx + by * by;
  const double E = p / kGammaM1 + 0.5 * rho * v_sq + 0.5 * b_sq;
  a.f[0][o] = (rho < 1e-8) ? 1e-8 : rho;   // the density floor of an upload (Grid.hpp:115)
  a.f[1][o] = rho * vx;
  a.f[2][o] = rho * vy;
  a.f[3][o] = E;
  a.f[4][o] = bx;
  a.f[5][o] = by;
}

// setups.fourier_ensemble_member: rho = 1, p = 1, v = sum of sines at the
// cell centres, B = (1, 0.5) + curl of the corner potential
// Az = sum (a / |k|) cos(...) + 1.0 y - 0.5 x, face B from its differences.
__global__ __launch_bounds__(256) void k_init_fourier(InitArgs a, const FourierMode* __restrict__ modes,
                                                      int nmodes) {
#pragma clang fp contract(off)
  const long long n = (long long)a.nx * a.ny;
  const double L = a.nx * a.dx;
  for (long long o = blockIdx.x * (long long)blockDim.x + threadIdx.x; o < n; o += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(o % a.nx), j = (int)(o / a.nx);
    const double X = (i + 0.5) * a.dx, Y = (j + 0.5) * a.dy;
    const double XC = i * a.dx, YC = j * a.dy, XE = (i + 1) * a.dx, YN = (j + 1) * a.dy;
    double vx = 0.0, vy = 0.0, az = 0.0, az_n = 0.0, az_e = 0.0;
    for (int m = 0; m < nmodes; ++m) {
      const FourierMode& md = modes[m];
      const double arg = kTwoPi * (md.kx * X + md.ky * Y) / L;
      vx += md.a_vx * sin(arg + md.ph_vx);
      vy += md.a_vy * sin(arg + md.ph_vy);
      const double k2 = (kTwoPi / L) * sqrt(md.kx * md.kx + md.ky * md.ky);
      const double c = md.a_az / k2;
      az += c * cos(kTwoPi * (md.kx * XC + md.ky * YC) / L + md.ph_az);
      az_n += c * cos(kTwoPi * (md.kx * XC + md.ky * YN) / L + md.ph_az);
      az_e += c * cos(kTwoPi * (md.kx * XE + md.ky * YC) / L + md.ph_az);
    }
    az += 1.0 * YC - 0.5 * XC;
    az_n += 1.0 * YN - 0.5 * XC;
    az_e += 1.0 * YC - 0.5 * XE;
    const double bx = (az_n - az) / a.dy;
    const double by = -(az_e - az) / a.dx;
    store_state(a, o, 1.0, vx, vy, 1.0, bx, by);
  }
}

// Core profile G(s) = exp((cos(2 pi s) - 1) / (2 pi^2 R^2)), R = 0.2.
__device__ __forceinline__ double column_profile(double s) {
#pragma clang fp contract(off)
  const double den = 2.0 * (3.141592653589793 * 3.141592653589793) * (0.2 * 0.2);
  return exp((cos(kTwoPi * s) - 1.0) / den);
}

// Background potential 0.02 sin(2 pi x + 0.5) sin(2 pi y + 0.9) / (2 pi).
__device__ __forceinline__ double column_potential(double xc, double yc) {
#pragma clang fp contract(off)
  return 0.02 * sin(kTwoPi * xc + 0.5) * sin(kTwoPi * yc + 0.9) / kTwoPi;
}

// setups.kink_column / sausage_column on [0, 1)^2: the displaced core
// (xi(y) = sum_k a_k sin(2 pi k y + phi_k)), the axial field B0 G, pressure
// balance, the weak smooth background and the drift (0.3, 0.2).
__global__ __launch_bounds__(256) void k_init_column(InitArgs a, const double* __restrict__ amps,
                                                     const double* __restrict__ phases, int nmodes, int sausage) {
#pragma clang fp contract(off)
  const long long n = (long long)a.nx * a.ny;
  const double dx = a.dx;
  for (long long o = blockIdx.x * (long long)blockDim.x + threadIdx.x; o < n; o += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(o % a.nx), j = (int)(o / a.nx);
    const double X = (i + 0.5) * dx, Y = (j + 0.5) * dx;
    const double XC = i * dx, YC = j * dx, XE = (i + 1) * dx, YN = (j + 1) * dx;
    double xi = 0.0;
    for (int k = 0; k < nmodes; ++k) xi += amps[k] * sin(kTwoPi * (Y * (double)(k + 1)) + phases[k]);
    double s;
    if (sausage) {
      const double w = -sin(kTwoPi * (X - 0.5)) / kTwoPi;
      s = X - 0.5 + xi * w;
    } else {
      s = X - 0.5 - xi;
    }
    const double g = column_profile(s);
    const double gb = column_profile(X - 0.5);
    const double vx = 0.3 + 0.05 * sin(kTwoPi * Y + 0.3) + 0.03 * cos(2.0 * kTwoPi * Y + kTwoPi * X + 1.0);
    const double vy = 0.2 + 0.05 * sin(kTwoPi * X + 0.7) + 0.02 * cos(kTwoPi * (X - Y) + 0.4);
    const double mr = 1.0 + 0.05 * sin(kTwoPi * (X + 2.0 * Y) + 0.2);
    const double mp = 1.0 + 0.05 * cos(kTwoPi * (2.0 * X - Y) + 1.3);
    const double az = column_potential(XC, YC);
    const double bx = (column_potential(XC, YN) - az) / dx;
    const double byb = -(column_potential(XE, YC) - az) / dx;
    const double rho = (0.2 + g) * mr;
    const double p = (0.1 + 0.5 * 25.0 * (1.0 - g * g)) * mp;
    const double by = 5.0 * gb + byb;
    store_state(a, o, rho, vx, vy, p, bx, by);
  }
}

// The reference's problem setups (include/benchmarks.hpp), as the package's
// host restatement (setups.py) evaluates them on the reference's 2 pi domain:
// cell centres X = (i + 0.5) dx, corners XC = i dx of the compact cell (i, j).
constexpr double kPiRef = 3.14159265359;            // the reference's truncated pi (benchmarks.hpp)
constexpr double kGamma = 5.0 / 3.0;

// orszag_tang_initial_conditions, benchmarks.hpp:15-76
__device__ __forceinline__ double ot_az(double xc, double yc) {
#pragma clang fp contract(off)
  return 0.5 * cos(2 * xc) + cos(yc);
}

__global__ __launch_bounds__(256) void k_init_reference(InitArgs a, int setup) {
#pragma clang fp contract(off)
  const long long n = (long long)a.nx * a.ny;
  const double dx = a.dx, dy = a.dy;
  const int nx = a.nx, ny = a.ny;
  for (long long o = blockIdx.x * (long long)blockDim.x + threadIdx.x; o < n; o += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(o % nx), j = (int)(o / nx);
    const double X = (i + 0.5) * dx, Y = (j + 0.5) * dy;
    const double XC = i * dx, YC = j * dy, XE = (i + 1) * dx, YN = (j + 1) * dy;
    double rho = 1.0, vx = 0.0, vy = 0.0, p = 1.0, bx = 0.0, by = 0.0;
    switch (setup) {
      case MHD_SETUP_ORSZAG_TANG: {
        const double az = ot_az(XC, YC);
        bx = (ot_az(XC, YN) - az) / dy;
        by = -(ot_az(XE, YC) - az) / dx;
        rho = 25.0 / 9.0;
        vx = -sin(Y);
        vy = sin(X);
        p = 5.0 / 3.0;
        break;
      }
      case MHD_SETUP_DECAYING_SINE: {   // decaying_magnetic_sine_wave_initial_conditions, benchmarks.hpp:352-417
        const double kx = 2 * kPiRef / (nx * dx), ky = 2 * kPiRef / (ny * dy);
        auto az = [&](double xc, double yc) -> double { return (1.0 / kx) * cos(kx * xc) * sin(ky * yc); };
        const double a0 = az(XC, YC);
        bx = (az(XC, YN) - a0) / dy;
        by = -(az(XE, YC) - a0) / dx;
        break;
      }
      case MHD_SETUP_ALFVEN: {          // alfven_wave_test_initial_values, benchmarks.hpp:78-126
        const double kx = 2.0 * kPiRef / (nx * dx);
        by = 1e-3 * sin(kx * X);
        vy = -by / sqrt(1.0);
        bx = 1.0;
        break;
      }
      case MHD_SETUP_GAUSSIAN: {        // gaussian_density_pulse_test_initial_values, benchmarks.hpp:128-176
        const double x_mid = 0.5 * nx * dx, y_mid = 0.5 * nx * dy;   // y_mid from nx, as in the reference
        const double d2 = (X - x_mid) * (X - x_mid) + (Y - y_mid) * (Y - y_mid);
        rho = 1.0 + exp(-5 * d2 / 4.0);
        vx = vy = 1.0;
        break;
      }
      case MHD_SETUP_ACOUSTIC: {        // acoustic_wave_test_initial_values, benchmarks.hpp:178-231
        const double x = i * dx;       // the reference evaluates this setup at the cell's left corner
        const double W = dx * nx;
        const double p_bg = 1.0 / kGamma;
        const double cs = sqrt(kGamma * p_bg / 1.0);
        const double xs = 1e-4 * sin(4 * kPiRef * x / W);
        rho = 1.0 + xs + 0.0;
        vx = (cs / 1.0) * xs;
        p = p_bg + (cs * cs) * (xs + 0.0);
        break;
      }
      case MHD_SETUP_SAUSAGE: {         // sausage_instability_initial_conditions, benchmarks.hpp:233-293
        const double k = 2.0 * kPiRef / (ny * dy);
        const double x_mid = 0.5 * (nx + 4) * dx;   // from the padded width, as in the reference
        const double r = X - x_mid;
        const double pert = 1.0 - 0.15 * cos(k * Y);
        rho = 1.0 * exp(-(r * r) / (0.04 * pert)) + 0.1;
        rho = (rho < 1e-4) ? 1e-4 : rho;
        p = 0.1 * exp(-(r * r) / (0.04 * pert)) + 0.1;
        p = (p < 1e-4) ? 1e-4 : p;
        by = 5.0 * (r / sqrt(0.04)) * exp(-(r * r) / 0.04);
        break;
      }
      case MHD_SETUP_KINK: {            // kink_instability_initial_conditions, benchmarks.hpp:295-350
        const double r = X - 0.5 * (nx + 4) * dx;
        const double env = exp(-(r * r) / 0.04);
        rho = env + 0.1;
        vx = 0.05 * sin(2.0 * kPiRef * Y) * env;
        p = env + 0.1;
        by = env;
        break;
      }
      default:
        break;
    }
    store_state(a, o, rho, vx, vy, p, bx, by);
  }
}

int init_blocks(long long n) { return (int)std::min<long long>((n + 255) / 256, 8192); }

}  // namespace

}  // namespace mhd

using namespace mhd;

namespace {

// Host checks shared by the batch and context entry points.
std::string check_square(int nx, int ny, double dx, double dy) {
  if (nx != ny || dx != dy) return "the ensemble setups are square (nx == ny, dx == dy)";
  return "";
}

std::string check_fourier(int nmodes, const double* modes) {
  if (nmodes < 0 || nmodes > 4096) return "nmodes must be in [0, 4096]";
  if (nmodes > 0 && !modes) return "NULL mode table";
  for (int m = 0; m < nmodes; ++m) {
    const double kx = modes[8 * m], ky = modes[8 * m + 1];
    if (!(kx * kx + ky * ky > 0.0)) return "mode " + std::to_string(m) + ": zero wave vector";
  }
  return "";
}

std::string check_column(int nmodes, const double* amps, const double* phases) {
  if (nmodes < 0 || nmodes > 4096) return "nmodes must be in [0, 4096]";
  if (nmodes > 0 && (!amps || !phases)) return "NULL amplitudes or phases";
  return "";
}

// Copies the parameter table to the device, runs the generator on `stream`
// and waits for it (the table is freed afterwards).
template <typename Launch>
hipError_t run_generator(hipStream_t stream, const void* table, size_t bytes, Launch&& launch) {
  void* d = nullptr;
  hipError_t e = hipSuccess;
  if (bytes > 0) {
    if ((e = hipMalloc(&d, bytes)) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(d, table, bytes, hipMemcpyHostToDevice, stream)) != hipSuccess) {
      (void)hipFree(d);
      return e;
    }
  }
  launch(d);
  e = hipGetLastError();
  const hipError_t es = hipStreamSynchronize(stream);
  if (d) (void)hipFree(d);
  return e != hipSuccess ? e : es;
}

}  // namespace

extern "C" {

int mhd_batch_init_fourier(mhd_batch* b, int run, int nmodes, const double* modes) {
  if (!b) return MHD_EINVAL;
  if (run < 0 || run >= b->nruns) return mhd_batch_host::bfail(b, MHD_EINVAL, "mhd_batch_init_fourier: bad run index");
  b->cfl_fresh = false;
  const RunGeom& g = b->geom_h[run];
  std::string why = check_square(g.nx, g.ny, g.dx, g.dy);
  if (why.empty()) why = check_fourier(nmodes, modes);
  if (!why.empty()) return mhd_batch_host::bfail(b, MHD_EINVAL, "mhd_batch_init_fourier: " + why);
  InitArgs a{};
  for (int q = 0; q < 6; ++q) a.f[q] = b->U[q] + g.base;
  a.nx = g.nx;
  a.ny = g.ny;
  a.dx = g.dx;
  a.dy = g.dy;
  const long long n = (long long)g.nx * g.ny;
  hipError_t e = run_generator(b->stream, modes, sizeof(FourierMode) * (size_t)nmodes, [&](void* d) {
    hipLaunchKernelGGL(k_init_fourier, dim3(init_blocks(n)), dim3(256), 0, b->stream, a,
                       (const FourierMode*)d, nmodes);
  });
  if (e != hipSuccess) return mhd_batch_host::bfail(b, MHD_EHIP, std::string("mhd_batch_init_fourier: ") + hipGetErrorString(e));
  RunState z{};   // a fresh state restarts the run's driver (as mhd_batch_upload)
  e = hipMemcpy(b->rs + run, &z, sizeof(RunState), hipMemcpyHostToDevice);
  if (e != hipSuccess) return mhd_batch_host::bfail(b, MHD_EHIP, std::string("mhd_batch_init_fourier: ") + hipGetErrorString(e));
  // the run's first CFL step reads the primitives of the new state
  hipLaunchKernelGGL(k_raw_reset, dim3(init_blocks(n)), dim3(256), 0, b->stream, (const double*)b->U[0] + g.base,
                     b->rho_raw + g.base, (int)n);
  e = hipGetLastError();
  if (e != hipSuccess) return mhd_batch_host::bfail(b, MHD_EHIP, std::string("mhd_batch_init_fourier: ") + hipGetErrorString(e));
  return MHD_OK;
}

int mhd_batch_init_column(mhd_batch* b, int run, int sausage, int nmodes, const double* amps, const double* phases) {
  if (!b) return MHD_EINVAL;
  if (run < 0 || run >= b->nruns) return mhd_batch_host::bfail(b, MHD_EINVAL, "mhd_batch_init_column: bad run index");
  b->cfl_fresh = false;
  const RunGeom& g = b->geom_h[run];
  std::string why = check_square(g.nx, g.ny, g.dx, g.dy);
  if (why.empty()) why = check_column(nmodes, amps, phases);
  if (!why.empty()) return mhd_batch_host::bfail(b, MHD_EINVAL, "mhd_batch_init_column: " + why);
  InitArgs a{};
  for (int q = 0; q < 6; ++q) a.f[q] = b->U[q] + g.base;
  a.nx = g.nx;
  a.ny = g.ny;
  a.dx = g.dx;
  a.dy = g.dy;
  std::vector<double> table(2 * (size_t)nmodes);
  std::copy(amps, amps + nmodes, table.begin());
  std::copy(phases, phases + nmodes, table.begin() + nmodes);
  const long long n = (long long)g.nx * g.ny;
  hipError_t e = run_generator(b->stream, table.data(), sizeof(double) * table.size(), [&](void* d) {
    const double* t = (const double*)d;
    hipLaunchKernelGGL(k_init_column, dim3(init_blocks(n)), dim3(256), 0, b->stream, a, t, t ? t + nmodes : t,
                       nmodes, sausage ? 1 : 0);
  });
  if (e != hipSuccess) return mhd_batch_host::bfail(b, MHD_EHIP, std::string("mhd_batch_init_column: ") + hipGetErrorString(e));
  RunState z{};
  e = hipMemcpy(b->rs + run, &z, sizeof(RunState), hipMemcpyHostToDevice);
  if (e != hipSuccess) return mhd_batch_host::bfail(b, MHD_EHIP, std::string("mhd_batch_init_column: ") + hipGetErrorString(e));
  // the run's first CFL step reads the primitives of the new state
  hipLaunchKernelGGL(k_raw_reset, dim3(init_blocks(n)), dim3(256), 0, b->stream, (const double*)b->U[0] + g.base,
                     b->rho_raw + g.base, (int)n);
  e = hipGetLastError();
  if (e != hipSuccess) return mhd_batch_host::bfail(b, MHD_EHIP, std::string("mhd_batch_init_column: ") + hipGetErrorString(e));
  return MHD_OK;
}

int mhd_batch_init_setup(mhd_batch* b, int run, int setup) {
  if (!b) return MHD_EINVAL;
  if (run < 0 || run >= b->nruns) return mhd_batch_host::bfail(b, MHD_EINVAL, "mhd_batch_init_setup: bad run index");
  if (setup < MHD_SETUP_ORSZAG_TANG || setup > MHD_SETUP_KINK)
    return mhd_batch_host::bfail(b, MHD_EINVAL, "mhd_batch_init_setup: unknown setup");
  b->cfl_fresh = false;
  const RunGeom& g = b->geom_h[run];
  InitArgs a{};
  for (int q = 0; q < 6; ++q) a.f[q] = b->U[q] + g.base;
  a.nx = g.nx;
  a.ny = g.ny;
  a.dx = g.dx;
  a.dy = g.dy;
  const long long n = (long long)g.nx * g.ny;
  hipError_t e = run_generator(b->stream, nullptr, 0, [&](void*) {
    hipLaunchKernelGGL(k_init_reference, dim3(init_blocks(n)), dim3(256), 0, b->stream, a, setup);
  });
  if (e != hipSuccess) return mhd_batch_host::bfail(b, MHD_EHIP, std::string("mhd_batch_init_setup: ") + hipGetErrorString(e));
  RunState z{};
  e = hipMemcpy(b->rs + run, &z, sizeof(RunState), hipMemcpyHostToDevice);
  if (e != hipSuccess) return mhd_batch_host::bfail(b, MHD_EHIP, std::string("mhd_batch_init_setup: ") + hipGetErrorString(e));
  // the run's first CFL step reads the primitives of the new state
  hipLaunchKernelGGL(k_raw_reset, dim3(init_blocks(n)), dim3(256), 0, b->stream, (const double*)b->U[0] + g.base,
                     b->rho_raw + g.base, (int)n);
  e = hipGetLastError();
  if (e != hipSuccess) return mhd_batch_host::bfail(b, MHD_EHIP, std::string("mhd_batch_init_setup: ") + hipGetErrorString(e));
  return MHD_OK;
}

int mhd_init_setup(mhd_ctx* c, int setup) {
  if (!c) return MHD_EINVAL;
  if (setup < MHD_SETUP_ORSZAG_TANG || setup > MHD_SETUP_KINK)
    return mhd_host::fail(c, MHD_EINVAL, "mhd_init_setup: unknown setup");
  InitArgs a{};
  for (int q = 0; q < 6; ++q) a.f[q] = c->U[q];
  a.nx = c->nx;
  a.ny = c->ny;
  a.dx = c->dx;
  a.dy = c->dy;
  hipError_t e = run_generator(c->stream, nullptr, 0, [&](void*) {
    hipLaunchKernelGGL(k_init_reference, dim3(init_blocks(c->n)), dim3(256), 0, c->stream, a, setup);
  });
  if (e != hipSuccess) return mhd_host::fail(c, MHD_EHIP, std::string("mhd_init_setup: ") + hipGetErrorString(e));
  return mhd_host::reset_raw_density(c);   // the first CFL step reads the primitives of the new state
}

int mhd_init_fourier(mhd_ctx* c, int nmodes, const double* modes) {
  if (!c) return MHD_EINVAL;
  std::string why = check_square(c->nx, c->ny, c->dx, c->dy);
  if (why.empty()) why = check_fourier(nmodes, modes);
  if (!why.empty()) return mhd_host::fail(c, MHD_EINVAL, "mhd_init_fourier: " + why);
  InitArgs a{};
  for (int q = 0; q < 6; ++q) a.f[q] = c->U[q];
  a.nx = c->nx;
  a.ny = c->ny;
  a.dx = c->dx;
  a.dy = c->dy;
  hipError_t e = run_generator(c->stream, modes, sizeof(FourierMode) * (size_t)nmodes, [&](void* d) {
    hipLaunchKernelGGL(k_init_fourier, dim3(init_blocks(c->n)), dim3(256), 0, c->stream, a,
                       (const FourierMode*)d, nmodes);
  });
  if (e != hipSuccess) return mhd_host::fail(c, MHD_EHIP, std::string("mhd_init_fourier: ") + hipGetErrorString(e));
  return mhd_host::reset_raw_density(c);   // the first CFL step reads the primitives of the new state
}

int mhd_init_column(mhd_ctx* c, int sausage, int nmodes, const double* amps, const double* phases) {
  if (!c) return MHD_EINVAL;
  std::string why = check_square(c->nx, c->ny, c->dx, c->dy);
  if (why.empty()) why = check_column(nmodes, amps, phases);
  if (!why.empty()) return mhd_host::fail(c, MHD_EINVAL, "mhd_init_column: " + why);
  InitArgs a{};
  for (int q = 0; q < 6; ++q) a.f[q] = c->U[q];
  a.nx = c->nx;
  a.ny = c->ny;
  a.dx = c->dx;
  a.dy = c->dy;
  std::vector<double> table(2 * (size_t)nmodes);
  std::copy(amps, amps + nmodes, table.begin());
  std::copy(phases, phases + nmodes, table.begin() + nmodes);
  hipError_t e = run_generator(c->stream, table.data(), sizeof(double) * table.size(), [&](void* d) {
    const double* t = (const double*)d;
    hipLaunchKernelGGL(k_init_column, dim3(init_blocks(c->n)), dim3(256), 0, c->stream, a, t, t ? t + nmodes : t,
                       nmodes, sausage ? 1 : 0);
  });
  if (e != hipSuccess) return mhd_host::fail(c, MHD_EHIP, std::string("mhd_init_column: ") + hipGetErrorString(e));
  return mhd_host::reset_raw_density(c);   // the first CFL step reads the primitives of the new state
}

}  // extern "C"
