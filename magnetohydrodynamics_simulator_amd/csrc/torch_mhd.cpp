// torch_mhd.cpp -- LibTorch autograd Function over the C ABI's adjoint
// (include/mhd_mi355x.h): n IMEX steps of constant dt as a differentiable
// operator on float64 GPU tensors.
//
//   (rho, mx, my, E, bxf, byf) at step n = ImexSteps(fields at step 0, eta;
//                                           dx, dy, dt, n, checkpoint_every)
//
// Fields are compact (ny, nx) tensors in the device layout of the library
// (row-major, bxf on the left face, byf on the bottom face of each cell);
// eta is a 0-d float64 tensor.  forward creates a context, runs
// mhd_adjoint_forward (checkpoints every k steps, the implicit solutions
// kept) and returns the final fields; backward runs mhd_adjoint_backward on
// the incoming gradients and returns dL/d(fields) and dL/d eta.  The context
// lives as long as the autograd graph that holds it.  Tensors are exchanged
// by device pointer: the library runs on the HIP runtime torch loaded (both
// bind libamdhip64.so.7), and each call synchronises the device before the
// library reads torch's tensors; the library's entry points return after their
// own work has completed.
#include <torch/cuda.h>
#include <torch/extension.h>

#include <stdexcept>
#include <string>
#include <vector>

#include "mhd_mi355x.h"

namespace {

using torch::Tensor;
using torch::autograd::AutogradContext;
using torch::autograd::variable_list;

void check(int rc, mhd_ctx* c, const char* what) {
  if (rc != MHD_OK) throw std::runtime_error(std::string(what) + ": " + mhd_last_error(c));
}

// A one-byte CPU tensor that owns the context: its deleter destroys it.
Tensor owner_of(mhd_ctx* c) {
  return torch::from_blob(
      c, {1}, [](void* p) { mhd_destroy(static_cast<mhd_ctx*>(p)); }, torch::TensorOptions().dtype(torch::kUInt8));
}

mhd_ctx* ctx_of(const Tensor& owner) { return static_cast<mhd_ctx*>(owner.data_ptr()); }

Tensor field(const Tensor& t, const char* name, int64_t ny, int64_t nx, const torch::Device& dev) {
  TORCH_CHECK(t.dim() == 2 && t.size(0) == ny && t.size(1) == nx, name, ": expected a (", ny, ", ", nx,
              ") tensor, got ", t.sizes());
  TORCH_CHECK(t.scalar_type() == torch::kFloat64, name, ": expected float64");
  TORCH_CHECK(t.device() == dev, name, ": all fields must be on the same device");
  return t.contiguous();
}

struct ImexSteps : public torch::autograd::Function<ImexSteps> {
  static variable_list forward(AutogradContext* actx, Tensor rho, Tensor mx, Tensor my, Tensor E, Tensor bxf,
                               Tensor byf, Tensor eta, double dx, double dy, double dt, int64_t nsteps,
                               int64_t every, double cg_tol, int64_t cg_max_iter) {
    TORCH_CHECK(rho.is_cuda(), "ImexSteps: fields must be GPU tensors");
    TORCH_CHECK(eta.numel() == 1, "ImexSteps: eta must hold one value");
    const torch::Device dev = rho.device();
    const int64_t ny = rho.size(0), nx = rho.dim() == 2 ? rho.size(1) : -1;
    const Tensor in[6] = {field(rho, "rho", ny, nx, dev), field(mx, "mx", ny, nx, dev),
                          field(my, "my", ny, nx, dev),   field(E, "E", ny, nx, dev),
                          field(bxf, "bxf", ny, nx, dev), field(byf, "byf", ny, nx, dev)};
    mhd_opts o;
    mhd_default_opts(&o);
    o.cg_tol = cg_tol;
    o.cg_max_iter = cg_max_iter;
    mhd_ctx* c = nullptr;
    check(mhd_create((int)nx, (int)ny, dx, dy, eta.item<double>(), &o, &c), nullptr, "mhd_create");
    Tensor owner = owner_of(c);
    torch::cuda::synchronize(dev.index());
    const double* src[6];
    for (int q = 0; q < 6; ++q) src[q] = in[q].data_ptr<double>();
    check(mhd_upload_device(c, src), c, "mhd_upload_device");
    check(mhd_adjoint_forward(c, dt, (int)nsteps, (int)every), c, "mhd_adjoint_forward");
    variable_list out;
    double* dst[6];
    for (int q = 0; q < 6; ++q) {
      out.push_back(torch::empty_like(in[q]));
      dst[q] = out.back().data_ptr<double>();
    }
    check(mhd_download_device(c, dst), c, "mhd_download_device");
    actx->saved_data["owner"] = owner;
    actx->saved_data["eta_options_device"] = eta.device().str();
    return out;
  }

  static variable_list backward(AutogradContext* actx, variable_list grads) {
    mhd_ctx* c = ctx_of(actx->saved_data["owner"].toTensor());
    std::vector<Tensor> g(6);
    for (int q = 0; q < 6; ++q) g[q] = grads[q].defined() ? grads[q].contiguous() : Tensor();
    // undefined gradients (outputs that do not enter the loss) are zero
    Tensor proto;
    for (const Tensor& t : g)
      if (t.defined()) proto = t;
    TORCH_CHECK(proto.defined(), "ImexSteps.backward: no gradient");
    for (Tensor& t : g)
      if (!t.defined()) t = torch::zeros_like(proto);
    torch::cuda::synchronize(proto.device().index());
    const double* lam[6];
    double* lam0[6];
    variable_list out;
    for (int q = 0; q < 6; ++q) {
      lam[q] = g[q].data_ptr<double>();
      out.push_back(torch::empty_like(g[q]));
      lam0[q] = out.back().data_ptr<double>();
    }
    double eta_bar = 0.0;
    check(mhd_adjoint_backward_device(c, lam, lam0, &eta_bar), c, "mhd_adjoint_backward_device");
    const torch::Device eta_dev(actx->saved_data["eta_options_device"].toStringRef());
    out.push_back(torch::full({}, eta_bar, torch::TensorOptions().dtype(torch::kFloat64).device(eta_dev)));
    // no gradient for dx, dy, dt, nsteps, every, cg_tol, cg_max_iter
    for (int k = 0; k < 7; ++k) out.push_back(Tensor());
    return out;
  }
};

variable_list imex_steps(Tensor rho, Tensor mx, Tensor my, Tensor E, Tensor bxf, Tensor byf, Tensor eta, double dx,
                         double dy, double dt, int64_t nsteps, int64_t every, double cg_tol, int64_t cg_max_iter) {
  return ImexSteps::apply(rho, mx, my, E, bxf, byf, eta, dx, dy, dt, nsteps, every, cg_tol, cg_max_iter);
}

}  // namespace

PYBIND11_MODULE(TORCH_EXTENSION_NAME, m) {
  m.doc() = "LibTorch autograd Function of n IMEX steps of the MI355X MHD path (adjoint through the C ABI)";
  m.def("imex_steps", &imex_steps, "n IMEX steps of constant dt, differentiable in the fields and eta",
        py::arg("rho"), py::arg("mx"), py::arg("my"), py::arg("E"), py::arg("bxf"), py::arg("byf"), py::arg("eta"),
        py::arg("dx"), py::arg("dy"), py::arg("dt"), py::arg("nsteps"), py::arg("checkpoint_every") = 10,
        py::arg("cg_tol") = 2.220446049250313e-16, py::arg("cg_max_iter") = 0);
}
