// mhd_snapshots.hip -- mhd_run_resident_files: the driver loop of
// mhd_run_resident with the reference's snapshot files
// (write_data_to_binary_file, include/diagnostics.hpp:90-131, at the cadence
// of src/integrator_main.cpp:65-70, 102-111) written while the loop runs on.
//
// At a snapshot the compute stream forms the file's seven arrays (rho and E
// copied from the state, p, vx, vy, bxc, byc by the kernel mhd_download uses)
// and the div B partials into one of two device staging slots; a copy stream
// waits for them and moves the slot to pinned host memory; a writer thread
// waits for that copy, reduces the partials in mhd_divb's order, expands each
// compact array into the padded layout with mhd_download's ghost fill and
// writes the file.  The files are byte for byte those of mhd_write_snapshot
// called from a mhd_run_resident callback.  The compute stream never waits
// for the host: it waits (on the device) only for the copy out of a slot it
// is about to refill, and the host blocks only when the writer still holds
// the host slot the next copy needs (it has fallen two snapshots behind).
#include <hip/hip_runtime.h>

#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/mhd_mi355x.h"
#include "mhd_context.hpp"
#include "mhd_kernels.hpp"

using namespace mhd_host;
using mhd::DiagArgs;

namespace {

constexpr int kSlots = 2;
constexpr int kFields = 7;   // the file's order: rho, E, p, vx, vy, bxc, byc

struct Job {
  int slot;
  int frame;
  double dt;
};

struct AsyncFiles {
  mhd_ctx* c = nullptr;
  std::string fmt;
  int rate = 0;
  int device = 0;
  int nx = 0, ny = 0, open_x = 0, open_y = 0;
  double dx = 0.0, dy = 0.0;
  size_t n = 0;                    // cells
  int tiles = 0;                   // div B partials: 2 per tile
  size_t slot_doubles = 0;         // kFields n + 2 tiles
  double* dev[kSlots] = {};
  double* host[kSlots] = {};       // pinned
  hipStream_t copy = nullptr;
  hipEvent_t formed[kSlots] = {};
  hipEvent_t copied[kSlots] = {};
  bool recorded[kSlots] = {};      // copied[s] has been recorded (compute-thread state)
  int next = 0;                    // compute-thread state
  // shared with the writer thread, under mu
  std::mutex mu;
  std::condition_variable cv;
  std::deque<Job> queue;
  bool held[kSlots] = {};          // the writer holds host[s]
  bool stop = false;
  int err = MHD_OK;
  std::string err_msg;
  std::thread writer;

  void set_error_locked(int code, const std::string& m) {
    if (err == MHD_OK) {
      err = code;
      err_msg = m;
    }
  }

  // One file from host slot j.slot; returns MHD_OK or an error with msg.
  // Reads only what the compute thread does not change while the slot is held.
  int write_file(const Job& j, std::vector<double>& padded, std::string& msg) const {
    const size_t np = (size_t)(nx + 4) * (size_t)(ny + 4), nxt = (size_t)nx + 4;
    const double* h = host[j.slot];
    const double* part = h + kFields * n;
    double sum = 0.0, mx = 0.0;
    for (int b = 0; b < tiles; ++b) {   // mhd_divb's order
      sum += part[2 * b];
      if (part[2 * b + 1] > mx) mx = part[2 * b + 1];
    }
    const double rms = std::sqrt(sum / (double)(nx * ny));
    const double header[9] = {(double)nx, (double)ny, 2.0, (double)rate, dx, dy, j.dt, rms, mx};
    char path[4096];
    const int len = std::snprintf(path, sizeof(path), fmt.c_str(), j.frame);
    if (len < 0 || len >= (int)sizeof(path)) {
      msg = "mhd_run_resident_files: file name too long";
      return MHD_EINVAL;
    }
    std::FILE* f = std::fopen(path, "wb");
    if (!f) {
      msg = std::string("mhd_run_resident_files: cannot open ") + path;
      return MHD_EINVAL;
    }
    bool ok = std::fwrite(header, sizeof(double), 9, f) == 9;
    for (int q = 0; q < kFields && ok; ++q) {
      const double* src = h + q * n;
      for (int r = 0; r < ny; ++r)
        std::memcpy(padded.data() + (r + 2) * nxt + 2, src + (size_t)r * nx, (size_t)nx * sizeof(double));
      fill_ghosts(padded.data(), nx, ny, open_x, open_y);
      ok = std::fwrite(padded.data(), sizeof(double), np, f) == np;
    }
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) {
      msg = std::string("mhd_run_resident_files: write to ") + path + " failed";
      return MHD_EINVAL;
    }
    return MHD_OK;
  }

  void writer_loop() {
    (void)hipSetDevice(device);
    std::vector<double> padded((size_t)(nx + 4) * (size_t)(ny + 4));
    for (;;) {
      Job j;
      bool failed;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return stop || !queue.empty(); });
        if (queue.empty()) return;
        j = queue.front();
        queue.pop_front();
        failed = err != MHD_OK;
      }
      int code = MHD_OK;
      std::string msg;
      if (!failed) {
        const hipError_t e = hipEventSynchronize(copied[j.slot]);
        if (e != hipSuccess) {
          code = MHD_EHIP;
          msg = std::string("snapshot copy: ") + hipGetErrorString(e);
        } else {
          code = write_file(j, padded, msg);
        }
      }
      {
        std::lock_guard<std::mutex> lk(mu);
        if (code != MHD_OK) set_error_locked(code, msg);
        held[j.slot] = false;
      }
      cv.notify_all();
    }
  }

  // Allocation and the writer thread.
  int open(mhd_ctx* ctx, const char* path_format, int snapshot_rate) {
    c = ctx;
    fmt = path_format;
    rate = snapshot_rate;
    nx = c->nx;
    ny = c->ny;
    dx = c->dx;
    dy = c->dy;
    open_x = c->o.bc_x == MHD_BC_OUTFLOW;
    open_y = c->o.bc_y == MHD_BC_OUTFLOW;
    n = (size_t)c->n;
    tiles = c->tiles;
    slot_doubles = kFields * n + 2 * (size_t)tiles;
    if (hipGetDevice(&device) != hipSuccess) return fail(c, MHD_EHIP, "mhd_run_resident_files: no device");
    for (int s = 0; s < kSlots; ++s) {
      if (hipMalloc(&dev[s], slot_doubles * sizeof(double)) != hipSuccess) {
        dev[s] = nullptr;
        return fail(c, MHD_ENOMEM, "mhd_run_resident_files: staging allocation failed");
      }
      if (hipHostMalloc(&host[s], slot_doubles * sizeof(double)) != hipSuccess) {
        host[s] = nullptr;
        return fail(c, MHD_ENOMEM, "mhd_run_resident_files: pinned allocation failed");
      }
      HIP_OR_FAIL(c, hipEventCreateWithFlags(&formed[s], hipEventDisableTiming));
      HIP_OR_FAIL(c, hipEventCreateWithFlags(&copied[s], hipEventDisableTiming));
    }
    HIP_OR_FAIL(c, hipStreamCreateWithFlags(&copy, hipStreamNonBlocking));
    writer = std::thread([this] { writer_loop(); });
    return MHD_OK;
  }

  // Enqueues the snapshot of the context's current state (compute thread).
  int snapshot(int frame, double dt) {
    const int s = next;
    next ^= 1;
    {
      std::unique_lock<std::mutex> lk(mu);
      cv.wait(lk, [&] { return !held[s] || err != MHD_OK; });
      if (err != MHD_OK) return fail(c, err, err_msg);
      held[s] = true;
    }
    // the slot's previous copy must be done before the device refills it
    if (recorded[s]) HIP_OR_FAIL(c, hipStreamWaitEvent(c->stream, copied[s], 0));
    double* d = dev[s];
    double* dst[2] = {d, d + n};
    double* src[2] = {c->U[0], c->U[3]};
    int rc = copy_fields(c, dst, src, 2);
    if (rc) return rc;
    const DiagArgs dp = diag_args(c, false);
    rc = launch(c, 8, [&] {
      hipLaunchKernelGGL(mhd::k_prims, dim3(grid_1d(c->n)), dim3(256), 0, c->stream, dp, d + 3 * n, d + 4 * n,
                         d + 2 * n, d + 5 * n, d + 6 * n);
    });
    if (rc) return rc;
    DiagArgs dd = diag_args(c, false);
    dd.part = d + kFields * n;
    rc = launch_k(c, 8, mhd::k_divb, dim3(tiles), dim3(mhd::kTileX), dd);
    if (rc) return rc;
    HIP_OR_FAIL(c, hipEventRecord(formed[s], c->stream));
    HIP_OR_FAIL(c, hipStreamWaitEvent(copy, formed[s], 0));
    HIP_OR_FAIL(c, hipMemcpyAsync(host[s], d, slot_doubles * sizeof(double), hipMemcpyDeviceToHost, copy));
    HIP_OR_FAIL(c, hipEventRecord(copied[s], copy));
    recorded[s] = true;
    {
      std::lock_guard<std::mutex> lk(mu);
      queue.push_back(Job{s, frame, dt});
    }
    cv.notify_all();
    return MHD_OK;
  }

  // Waits for every file, joins the writer and frees everything; returns the
  // first writer error (MHD_OK if none).
  int close() {
    if (writer.joinable()) {
      {
        std::lock_guard<std::mutex> lk(mu);
        stop = true;
      }
      cv.notify_all();
      writer.join();
    }
    if (copy) {
      (void)hipStreamSynchronize(copy);
      (void)hipStreamDestroy(copy);
      copy = nullptr;
    }
    for (int s = 0; s < kSlots; ++s) {
      if (formed[s]) (void)hipEventDestroy(formed[s]);
      if (copied[s]) (void)hipEventDestroy(copied[s]);
      if (dev[s]) (void)hipFree(dev[s]);
      if (host[s]) (void)hipHostFree(host[s]);
      formed[s] = copied[s] = nullptr;
      dev[s] = host[s] = nullptr;
    }
    return err;
  }
};

int on_snapshot(mhd_ctx*, int frame, double, double dt, void* user) {
  return static_cast<AsyncFiles*>(user)->snapshot(frame, dt);
}

}  // namespace

extern "C" {

int mhd_run_resident_files(mhd_ctx* c, double t_end, double dt_override, int max_steps, int snapshot_every,
                           const char* path_format, int snapshot_rate, int* steps_out, double* t_out,
                           double* step_dt, int* cg_iters) {
  if (!c) return MHD_EINVAL;
  if (snapshot_every <= 0 || !path_format)
    return fail(c, MHD_EINVAL, "mhd_run_resident_files: need snapshot_every > 0 and a path format");
  {
    // exactly one integer conversion (the frame), no other conversions
    int conv = 0;
    for (const char* p = path_format; *p; ++p) {
      if (*p != '%') continue;
      if (p[1] == '%') {
        ++p;
        continue;
      }
      const char* q = p + 1;
      while (*q && std::strchr("0123456789-+ #", *q)) ++q;
      if (*q != 'd' && *q != 'i') return fail(c, MHD_EINVAL, "mhd_run_resident_files: path format takes one %d");
      ++conv;
      p = q;
    }
    if (conv != 1) return fail(c, MHD_EINVAL, "mhd_run_resident_files: path format takes one %d");
  }
  AsyncFiles A;
  int rc = A.open(c, path_format, snapshot_rate);
  if (rc == MHD_OK)
    rc = mhd_run_resident(c, t_end, dt_override, max_steps, snapshot_every, on_snapshot, &A, steps_out, t_out, step_dt,
                          cg_iters);
  const int werr = A.close();
  if (rc == MHD_OK && werr != MHD_OK) rc = fail(c, werr, A.err_msg);
  return rc;
}

}  // extern "C"
