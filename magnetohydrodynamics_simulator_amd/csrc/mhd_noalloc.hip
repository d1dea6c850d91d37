// mhd_noalloc.hip -- mhd_cfl_dt_noalloc: the CFL time step of mhd_cfl_dt
// (CFL_condition_checker, diagnostics.hpp:44-88) for callers that cannot
// allocate device memory (code under HIP graph capture, serving loops that
// forbid hipMalloc).  No context: the state is the caller's compact device
// arrays, the result goes to a device double, everything runs on the
// caller's stream with no host synchronisation.
//
// Scratch: the per-tile partials (max wave speed, non-finite count) of the
// CFL kernel live in the file-scope device array `scratch` below
// (2 * kNoAllocMaxTiles doubles = 64 KiB per device, held until the process
// ends).  Its device address is looked up once per device with
// hipGetSymbolAddress and kept; host code never passes the array's name.
// Every call on a device shares that device's array: calls that overlap on
// one device (different streams, or host threads) race on it.  Each call
// writes every partial it reads, so no clearing is needed.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <string>

#include "../../include/mhd_mi355x.h"
#include "mhd_context.hpp"
#include "mhd_kernels.hpp"

namespace {

constexpr int kNoAllocMaxTiles = 4096;
constexpr int kMaxDevices = 64;

__device__ __align__(256) double scratch[2 * kNoAllocMaxTiles];

double* g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mu;

// errors go where mhd_last_error(NULL) reads them
int noalloc_fail(int code, const std::string& msg) { return mhd_host::fail(nullptr, code, msg); }

// The device address of this device's scratch array (looked up on the first
// call on the device).
int scratch_ptr(double** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
    return noalloc_fail(MHD_EHIP, "mhd_cfl_dt_noalloc: no current HIP device");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return noalloc_fail(MHD_EHIP, std::string("mhd_cfl_dt_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return MHD_OK;
}

// dt from the nb partials, as mhd_cfl_dt forms it on the host (the same
// operations in the same order; a max and a count, exact in any grouping).
// A non-finite wave speed anywhere gives dt = NaN.
__global__ __launch_bounds__(64) void k_cfl_dt_from_partials(const double* part, int nb, double dx, double dy,
                                                              double eta, double cfl, double* dt_out) {
  __shared__ double lm[64], lb[64];
  double m = 1e-8, bad = 0.0;
  for (int b = threadIdx.x; b < nb; b += 64) {
    const double v = part[2 * b];
    if (v > m) m = v;
    bad += part[2 * b + 1];
  }
  lm[threadIdx.x] = m;
  lb[threadIdx.x] = bad;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int t = 1; t < 64; ++t) {
      if (lm[t] > m) m = lm[t];
      bad += lb[t];
    }
    const double h = (dy < dx) ? dy : dx;
    double dt_res = INFINITY;
    if (eta != 0.0) dt_res = h * h / (4.0 * eta);
    const double courant = cfl / 0.9;
    const double dt_hyp = courant * h / m;
    const double lim = (dt_res < dt_hyp) ? dt_res : dt_hyp;
    dt_out[0] = (bad > 0.0) ? NAN : lim * 0.9;
  }
}

}  // namespace

extern "C" {

int mhd_cfl_dt_noalloc(int nx, int ny, double dx, double dy, double eta, double cfl, int bc_x, int bc_y,
                       const double* const* fields, double* dt_device, void* stream) {
  if (nx < 4 || ny < 4 || (long long)nx * ny > (1LL << 30))
    return noalloc_fail(MHD_EINVAL, "mhd_cfl_dt_noalloc: need nx, ny >= 4 and nx * ny <= 2^30");
  if (!(dx > 0) || !(dy > 0) || !std::isfinite(dx) || !std::isfinite(dy) || !(eta >= 0) || !std::isfinite(eta) ||
      !(cfl > 0) || !std::isfinite(cfl))
    return noalloc_fail(MHD_EINVAL, "mhd_cfl_dt_noalloc: dx, dy, cfl must be positive and finite, eta >= 0");
  if ((bc_x != MHD_BC_PERIODIC && bc_x != MHD_BC_OUTFLOW) || (bc_y != MHD_BC_PERIODIC && bc_y != MHD_BC_OUTFLOW))
    return noalloc_fail(MHD_EINVAL, "mhd_cfl_dt_noalloc: bc_x / bc_y must be MHD_BC_PERIODIC or MHD_BC_OUTFLOW");
  if (!fields || !dt_device) return noalloc_fail(MHD_EINVAL, "mhd_cfl_dt_noalloc: NULL argument");
  for (int q = 0; q < 6; ++q)
    if (!fields[q]) return noalloc_fail(MHD_EINVAL, "mhd_cfl_dt_noalloc: NULL field");
  // the tiles of a context's CFL partials (mhd_create): kTileX columns x
  // fin_rows rows, about kFinTargetTiles of them
  const int tiles_x = (nx + mhd::kTileX - 1) / mhd::kTileX;
  int fin_y = mhd::kFinTargetTiles / tiles_x;
  if (fin_y < 1) fin_y = 1;
  int rows = (ny + fin_y - 1) / fin_y;
  if (rows < 4) rows = 4;
  const int tiles_y = (ny + rows - 1) / rows;
  const long long nb = (long long)tiles_x * tiles_y;
  if (nb > kNoAllocMaxTiles)
    return noalloc_fail(MHD_EINVAL, "mhd_cfl_dt_noalloc: the grid needs " + std::to_string(nb) +
                                        " partial tiles, more than the scratch holds (" +
                                        std::to_string(kNoAllocMaxTiles) + "; nx <= 256 * 4096)");
  double* part = nullptr;
  if (int rc = scratch_ptr(&part)) return rc;
  mhd::DiagArgs d{};
  d.nx = nx;
  d.ny = ny;
  d.ty = rows;
  d.tiles_y = tiles_y;
  d.open_x = bc_x == MHD_BC_OUTFLOW;
  d.open_y = bc_y == MHD_BC_OUTFLOW;
  d.dx = dx;
  d.dy = dy;
  for (int q = 0; q < 6; ++q) d.s[q] = fields[q];
  d.part = part;
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(mhd::k_cfl, dim3((unsigned)nb), dim3(mhd::kTileX), 0, s, d);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return noalloc_fail(MHD_EHIP, std::string("mhd_cfl_dt_noalloc: launch: ") + hipGetErrorString(e));
  hipLaunchKernelGGL(k_cfl_dt_from_partials, dim3(1), dim3(64), 0, s, (const double*)part, (int)nb, dx, dy, eta, cfl,
                     dt_device);
  e = hipGetLastError();
  if (e != hipSuccess) return noalloc_fail(MHD_EHIP, std::string("mhd_cfl_dt_noalloc: launch: ") + hipGetErrorString(e));
  return MHD_OK;
}

}  // extern "C"
