// PyTorch-ROCm binding of libdrot.so: the `reg_drot` module of the reference
// (pytorch/src/drot_cuda.cpp:26-62) with the same two functions, argument
// meanings and TORCH_CHECK error behaviour, layered on the C ABI of
// include/drot.h.  Differences from the reference, all additive:
//   * it runs on the caller's current HIP stream (the reference uses the legacy
//     default stream) under a device guard for c's device;
//   * it also checks dtype (the reference relies on data_ptr<float>() throwing);
//   * *_ex variants return (x, info) with n_iter, objective and residual, which
//     the reference computes and discards (qr_drot_cuda_kernel.cu:36,77).
#include <torch/extension.h>
#include <ATen/hip/impl/HIPGuardImplMasqueradingAsCUDA.h>
#include <ATen/hip/impl/HIPStreamMasqueradingAsCUDA.h>

#include "drot.h"

namespace {

#define CHECK_CUDA(x) TORCH_CHECK(x.device().is_cuda(), #x " must be a CUDA tensor")
#define CHECK_CONTIGUOUS(x) TORCH_CHECK(x.is_contiguous(), #x " must be contiguous")
#define CHECK_FLOAT(x) TORCH_CHECK(x.scalar_type() == torch::kFloat32, #x " must be float32")
#define CHECK_INPUT(x) \
    CHECK_CUDA(x);     \
    CHECK_CONTIGUOUS(x); \
    CHECK_FLOAT(x)

void check_shapes(const torch::Tensor &c, const torch::Tensor &p, const torch::Tensor &q) {
    CHECK_INPUT(c);
    CHECK_INPUT(p);
    CHECK_INPUT(q);
    TORCH_CHECK(c.dim() == 2, "C must be a 2-D tensor");
    TORCH_CHECK(p.device() == c.device() && q.device() == c.device(),
                "c, p and q must be on the same device");
    const int64_t n_rows = p.numel(), n_cols = q.numel();
    // pytorch/src/qr_drot_cuda_kernel.cu:92-93
    TORCH_CHECK(c.size(0) == n_cols, "C.size(0) does not match with q.numel()");
    TORCH_CHECK(c.size(1) == n_rows, "C.size(1) does not match with p.numel()");
    TORCH_CHECK(n_rows > 0 && n_cols > 0, "p and q must be non-empty");
    TORCH_CHECK(n_rows <= INT32_MAX && n_cols <= INT32_MAX, "p or q too long");
}

void check_status(drot_status s, const char *what) {
    TORCH_CHECK(s == DROT_OK, what, " failed: ", drot_status_string(s), ": ",
                drot_last_error());
}

py::dict info_dict(const drot_result &r) {
    py::dict d;
    d["n_iter"] = r.n_iter;
    d["converged"] = (bool)r.converged;
    d["iterations_run"] = r.iterations_run;
    d["objective"] = r.objective;
    d["objective_f64"] = r.objective_f64;
    d["residual"] = r.residual;
    return d;
}

std::pair<torch::Tensor, drot_result> qr_impl(torch::Tensor c, torch::Tensor p,
                                              torch::Tensor q, const drot_qr_params &prm) {
    const int n_rows = (int)p.numel(), n_cols = (int)q.numel();
    const c10::hip::HIPGuardMasqueradingAsCUDA guard(c.device());
    auto x = torch::empty_like(c);
    size_t ws_bytes = 0;
    check_status(drot_qr_workspace_size(n_rows, n_cols, &ws_bytes), "drot_qr_workspace_size");
    auto ws = torch::empty({(int64_t)ws_bytes}, c.options().dtype(torch::kUInt8));
    hipStream_t stream = c10::hip::getCurrentHIPStreamMasqueradingAsCUDA(c.device().index()).stream();
    drot_result r{};
    check_status(drot_qr_solve_f32(c.data_ptr<float>(), p.data_ptr<float>(),
                                   q.data_ptr<float>(), n_rows, n_cols, &prm,
                                   x.data_ptr<float>(), ws.data_ptr(), ws_bytes,
                                   (void *)stream, &r),
                 "drot_qr_solve_f32");
    return {x, r};
}

std::pair<torch::Tensor, drot_result> glr_impl(torch::Tensor c, torch::Tensor p,
                                               torch::Tensor q, const drot_glr_params &prm) {
    const int n_rows = (int)p.numel(), n_cols = (int)q.numel();
    const c10::hip::HIPGuardMasqueradingAsCUDA guard(c.device());
    auto x = torch::empty_like(c);
    size_t ws_bytes = 0;
    check_status(drot_glr_workspace_size(n_rows, n_cols, prm.n_groups, &ws_bytes),
                 "drot_glr_workspace_size");
    auto ws = torch::empty({(int64_t)ws_bytes}, c.options().dtype(torch::kUInt8));
    hipStream_t stream = c10::hip::getCurrentHIPStreamMasqueradingAsCUDA(c.device().index()).stream();
    drot_result r{};
    check_status(drot_glr_solve_f32(c.data_ptr<float>(), p.data_ptr<float>(),
                                    q.data_ptr<float>(), n_rows, n_cols, &prm,
                                    x.data_ptr<float>(), ws.data_ptr(), ws_bytes,
                                    (void *)stream, &r),
                 "drot_glr_solve_f32");
    return {x, r};
}

drot_qr_params qr_params(const torch::Tensor &p, const torch::Tensor &q, float rho,
                         float r_weight, int max_iters, float eps) {
    drot_qr_params prm;
    check_status(drot_qr_params_from_rho(rho, r_weight, (int)p.numel(), (int)q.numel(),
                                         max_iters, eps, &prm),
                 "drot_qr_params_from_rho");
    return prm;
}

drot_glr_params glr_params(const torch::Tensor &p, const torch::Tensor &q, int n_class,
                           float rho, float r_weight, int max_iters, float eps) {
    TORCH_CHECK(n_class >= 1 && n_class <= p.numel(), "n_class must be in [1, p.numel()]");
    drot_glr_params prm;
    check_status(drot_glr_params_from_rho(rho, r_weight, (int)p.numel(), (int)q.numel(),
                                          n_class, max_iters, eps, &prm),
                 "drot_glr_params_from_rho");
    return prm;
}

// reg_drot.quadratic_drot(c, p, q, rho, r_weight, max_iters, eps)
// (pytorch/src/drot_cuda.cpp:26-40)
torch::Tensor quadratic_drot(torch::Tensor c, torch::Tensor p, torch::Tensor q, float rho,
                             float r_weight, int max_iters, float eps) {
    check_shapes(c, p, q);
    return qr_impl(c, p, q, qr_params(p, q, rho, r_weight, max_iters, eps)).first;
}

py::tuple quadratic_drot_ex(torch::Tensor c, torch::Tensor p, torch::Tensor q, float rho,
                            float r_weight, int max_iters, float eps) {
    check_shapes(c, p, q);
    const drot_qr_params prm = qr_params(p, q, rho, r_weight, max_iters, eps);
    auto out = qr_impl(c, p, q, prm);
    py::dict d = info_dict(out.second);
    d["step_size"] = prm.step_size;
    d["scale"] = prm.scale;
    return py::make_tuple(out.first, d);
}

// dL/dC of a plan returned by quadratic_drot for the same (c, p, q, rho,
// r_weight): implicit differentiation on the plan's support (no counterpart in
// the reference, whose binding returns a tensor without gradient,
// pytorch/src/qr_drot_cuda_kernel.cu:99).  grad_obj: optional 0-d/1-element
// tensor on c's device, the gradient of the regularized objective.
py::tuple quadratic_drot_backward(torch::Tensor c, torch::Tensor p, torch::Tensor q,
                                  torch::Tensor x, torch::Tensor grad_x,
                                  c10::optional<torch::Tensor> grad_obj, float rho,
                                  float r_weight, float support_tol, int max_cg_iters,
                                  double cg_tol) {
    check_shapes(c, p, q);
    CHECK_INPUT(x);
    CHECK_INPUT(grad_x);
    TORCH_CHECK(x.sizes() == c.sizes() && grad_x.sizes() == c.sizes(),
                "x and grad_x must have the shape of C");
    TORCH_CHECK(x.device() == c.device() && grad_x.device() == c.device(),
                "x and grad_x must be on C's device");
    torch::Tensor gobj;
    if (grad_obj.has_value() && grad_obj->defined()) {
        gobj = grad_obj->to(c.device(), torch::kFloat32).reshape({-1}).contiguous();
        TORCH_CHECK(gobj.numel() == 1, "grad_obj must hold one value");
    }
    const int n_rows = (int)p.numel(), n_cols = (int)q.numel();
    const drot_qr_params prm = qr_params(p, q, rho, r_weight, 0, 0.0f);
    drot_backward_params bp{};
    bp.support_tol = support_tol;
    bp.max_cg_iters = max_cg_iters;
    bp.cg_tol = cg_tol;
    const c10::hip::HIPGuardMasqueradingAsCUDA guard(c.device());
    size_t ws_bytes = 0;
    check_status(drot_qr_backward_workspace_size(n_rows, n_cols, &prm, &bp, &ws_bytes),
                 "drot_qr_backward_workspace_size");
    auto ws = torch::empty({(int64_t)ws_bytes}, c.options().dtype(torch::kUInt8));
    auto grad_c = torch::empty_like(c);
    hipStream_t stream = c10::hip::getCurrentHIPStreamMasqueradingAsCUDA(c.device().index()).stream();
    drot_backward_info info{};
    check_status(drot_qr_backward_f32(c.data_ptr<float>(), x.data_ptr<float>(),
                                      grad_x.data_ptr<float>(),
                                      gobj.defined() ? gobj.data_ptr<float>() : nullptr, n_rows,
                                      n_cols, &prm, &bp, grad_c.data_ptr<float>(), ws.data_ptr(),
                                      ws_bytes, (void *)stream, &info),
                 "drot_qr_backward_f32");
    py::dict d;
    d["nnz"] = info.nnz;
    d["cg_iters"] = info.cg_iters;
    d["cg_converged"] = (bool)info.converged;
    d["cg_residual"] = info.cg_residual;
    return py::make_tuple(grad_c, d);
}

// reg_drot.group_lasso_drot(c, p, q, n_class, rho, r_weight, max_iters, eps)
// (pytorch/src/drot_cuda.cpp:42-57)
torch::Tensor group_lasso_drot(torch::Tensor c, torch::Tensor p, torch::Tensor q, int n_class,
                               float rho, float r_weight, int max_iters, float eps) {
    check_shapes(c, p, q);
    return glr_impl(c, p, q, glr_params(p, q, n_class, rho, r_weight, max_iters, eps)).first;
}

py::tuple group_lasso_drot_ex(torch::Tensor c, torch::Tensor p, torch::Tensor q, int n_class,
                              float rho, float r_weight, int max_iters, float eps) {
    check_shapes(c, p, q);
    auto out = glr_impl(c, p, q, glr_params(p, q, n_class, rho, r_weight, max_iters, eps));
    return py::make_tuple(out.first, info_dict(out.second));
}

}  // namespace

PYBIND11_MODULE(TORCH_EXTENSION_NAME, m) {
    m.def("quadratic_drot", &quadratic_drot, "Quadratic Regularized DROT");
    m.def("group_lasso_drot", &group_lasso_drot, "Group-Lasso Regularized DROT");
    m.def("quadratic_drot_ex", &quadratic_drot_ex,
          "Quadratic Regularized DROT, returning (x, info)");
    m.def("group_lasso_drot_ex", &group_lasso_drot_ex,
          "Group-Lasso Regularized DROT, returning (x, info)");
    m.def("quadratic_drot_backward", &quadratic_drot_backward,
          "Implicit-differentiation gradient dL/dC of a quadratic DROT plan, returning "
          "(grad_c, info)",
          py::arg("c"), py::arg("p"), py::arg("q"), py::arg("x"), py::arg("grad_x"),
          py::arg("grad_obj"), py::arg("rho"), py::arg("r_weight"), py::arg("support_tol") = 0.0f,
          py::arg("max_cg_iters") = 0, py::arg("cg_tol") = 0.0);
    m.attr("abi_version") = drot_abi_version();
}
