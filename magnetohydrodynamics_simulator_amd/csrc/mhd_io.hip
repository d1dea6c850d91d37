// mhd_io.hip -- snapshot and restart files of a context (C ABI mhd_write_snapshot,
// mhd_write_restart, mhd_read_restart).
//
// Snapshots follow write_data_to_binary_file (include/diagnostics.hpp:90-131
// of the reference), the format its reader scripts/display_functions.py:3-22
// expects: 9 doubles (nx, ny, ng, snapshot_rate, cell_width, cell_height,
// time_step, rms div B, max |div B|), then seven padded (ny+4) x (nx+4)
// arrays in the order rho, E, p, vx, vy, bxc, byc, ghosts included.  The
// primitives are those of fillGhostPeriodic + computePrimitives of the state
// (mhd_download).  The reference's computeCellB leaves the last padded row and
// column of bxc / byc untouched (Grid.hpp:82-97); here they hold the periodic
// ghost values (the reader drops all ghost layers).
//
// Restart files (the reference cannot resume): 8 bytes "MHDRST01", 7 doubles
// (nx, ny, ng, dx, dy, eta, t), then the six padded conserved arrays rho, mx,
// my, E, bxf, byf -- everything an exact restart of the step needs.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mhd_mi355x.h"
#include "mhd_context.hpp"

using namespace mhd_host;

namespace {

constexpr char kRestartMagic[8] = {'M', 'H', 'D', 'R', 'S', 'T', '0', '1'};

struct File {
  std::FILE* f = nullptr;
  ~File() {
    if (f) std::fclose(f);
  }
};

bool write_all(std::FILE* f, const void* p, size_t bytes) { return std::fwrite(p, 1, bytes, f) == bytes; }
bool read_all(std::FILE* f, void* p, size_t bytes) { return std::fread(p, 1, bytes, f) == bytes; }

}  // namespace

extern "C" {

int mhd_write_snapshot(mhd_ctx* c, const char* path, int snapshot_rate, double time_step) {
  if (!c) return MHD_EINVAL;
  if (!path) return fail(c, MHD_EINVAL, "mhd_write_snapshot: NULL path");
  const size_t np = (size_t)(c->nx + 4) * (size_t)(c->ny + 4);
  std::vector<double> rho(np), E(np), p(np), vx(np), vy(np), bxc(np), byc(np);
  int rc = mhd_download(c, rho.data(), nullptr, nullptr, E.data(), nullptr, nullptr, vx.data(), vy.data(), p.data(),
                        bxc.data(), byc.data());
  if (rc) return rc;
  double rms = 0.0, mx = 0.0;
  if ((rc = mhd_divb(c, &rms, &mx))) return rc;
  const double header[9] = {(double)c->nx, (double)c->ny, 2.0, (double)snapshot_rate, c->dx, c->dy, time_step, rms,
                            mx};
  File out;
  out.f = std::fopen(path, "wb");
  if (!out.f) return fail(c, MHD_EINVAL, std::string("mhd_write_snapshot: cannot open ") + path);
  bool ok = write_all(out.f, header, sizeof(header));
  for (const std::vector<double>* a : {&rho, &E, &p, &vx, &vy, &bxc, &byc})
    ok = ok && write_all(out.f, a->data(), np * sizeof(double));
  ok = (std::fclose(out.f) == 0) && ok;
  out.f = nullptr;
  if (!ok) return fail(c, MHD_EINVAL, std::string("mhd_write_snapshot: write to ") + path + " failed");
  return MHD_OK;
}

int mhd_write_restart(mhd_ctx* c, const char* path, double t) {
  if (!c) return MHD_EINVAL;
  if (!path) return fail(c, MHD_EINVAL, "mhd_write_restart: NULL path");
  const size_t np = (size_t)(c->nx + 4) * (size_t)(c->ny + 4);
  std::vector<double> f[6];
  for (auto& a : f) a.resize(np);
  int rc = mhd_download(c, f[0].data(), f[1].data(), f[2].data(), f[3].data(), f[4].data(), f[5].data(), nullptr,
                        nullptr, nullptr, nullptr, nullptr);
  if (rc) return rc;
  const double header[7] = {(double)c->nx, (double)c->ny, 2.0, c->dx, c->dy, c->eta, t};
  File out;
  out.f = std::fopen(path, "wb");
  if (!out.f) return fail(c, MHD_EINVAL, std::string("mhd_write_restart: cannot open ") + path);
  bool ok = write_all(out.f, kRestartMagic, sizeof(kRestartMagic)) && write_all(out.f, header, sizeof(header));
  for (const auto& a : f) ok = ok && write_all(out.f, a.data(), np * sizeof(double));
  ok = (std::fclose(out.f) == 0) && ok;
  out.f = nullptr;
  if (!ok) return fail(c, MHD_EINVAL, std::string("mhd_write_restart: write to ") + path + " failed");
  return MHD_OK;
}

int mhd_read_restart(mhd_ctx* c, const char* path, double* t) {
  if (!c) return MHD_EINVAL;
  if (!path) return fail(c, MHD_EINVAL, "mhd_read_restart: NULL path");
  File in;
  in.f = std::fopen(path, "rb");
  if (!in.f) return fail(c, MHD_EINVAL, std::string("mhd_read_restart: cannot open ") + path);
  char magic[8];
  double header[7];
  if (!read_all(in.f, magic, sizeof(magic)) || std::memcmp(magic, kRestartMagic, sizeof(magic)) != 0 ||
      !read_all(in.f, header, sizeof(header)))
    return fail(c, MHD_EINVAL, std::string("mhd_read_restart: ") + path + " is not a restart file");
  if (header[0] != (double)c->nx || header[1] != (double)c->ny || header[2] != 2.0 || header[3] != c->dx ||
      header[4] != c->dy)
    return fail(c, MHD_EINVAL, std::string("mhd_read_restart: ") + path + " holds another grid");
  const size_t np = (size_t)(c->nx + 4) * (size_t)(c->ny + 4);
  std::vector<double> f[6];
  for (auto& a : f) {
    a.resize(np);
    if (!read_all(in.f, a.data(), np * sizeof(double)))
      return fail(c, MHD_EINVAL, std::string("mhd_read_restart: ") + path + " is truncated");
  }
  int rc = mhd_upload(c, f[0].data(), f[1].data(), f[2].data(), f[3].data(), f[4].data(), f[5].data());
  if (rc) return rc;
  if (t) *t = header[6];
  return MHD_OK;
}

}  // extern "C"
