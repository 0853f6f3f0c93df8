// hipacc_autograd.cpp -- libtorch autograd Functions over the C ABI:
//   conv2d(x, w, boundary, const_val)       learnable float mask, Y = SUM_t w[t] X_bh(p+t-r)
//   bilateral(x, sigma_r, ksize, sigma_d, boundary, const_val)
//                                           bilateral filter with learnable range sigma
// Forward and backward run the gfx950 kernels of libhipacc_gfx950.so
// (hipacc_hip_local / hipacc_hip_local_grad, hipacc_hip_bilateral_train /
// hipacc_hip_bilateral_grad) on torch's current HIP stream.  Images are
// float32 (H, W) or (B, H, W) device tensors with contiguous rows.
#include <torch/extension.h>
#include <c10/hip/HIPStream.h>

#include <cstring>

#include "hipacc_hip.h"

namespace {

using torch::autograd::AutogradContext;
using torch::autograd::tensor_list;

void check(int status, const char* what) {
  TORCH_CHECK(status == HIPACC_OK, what, ": ", hipacc_hip_strerror(status), ": ",
              hipacc_hip_last_error());
}

void* current_stream() { return c10::hip::getCurrentHIPStream().stream(); }

void check_image(const torch::Tensor& t, const char* name) {
  TORCH_CHECK(t.is_cuda(), name, " must be a device tensor");
  TORCH_CHECK(t.scalar_type() == torch::kFloat32, name, " must be float32");
  TORCH_CHECK(t.dim() == 2 || t.dim() == 3, name, " must be (H, W) or (B, H, W)");
  TORCH_CHECK(t.stride(-1) == 1, name, " must have contiguous rows");
}

hipacc_view view_of(const torch::Tensor& t) {
  hipacc_view v;
  std::memset(&v, 0, sizeof(v));
  v.data = t.data_ptr();
  v.dtype = HIPACC_F32;
  v.stride = (int32_t)t.stride(-2);
  v.img_width = v.width = (int32_t)t.size(-1);
  v.img_height = v.height = (int32_t)t.size(-2);
  v.batch = t.dim() == 3 ? (int32_t)t.size(0) : 1;
  v.plane_stride = t.dim() == 3 ? t.stride(0) : 0;
  return v;
}

class LearnableConv : public torch::autograd::Function<LearnableConv> {
 public:
  static torch::Tensor forward(AutogradContext* ctx, torch::Tensor x, torch::Tensor w,
                               int64_t boundary, double const_val) {
    check_image(x, "x");
    TORCH_CHECK(w.is_cuda() && w.scalar_type() == torch::kFloat32 && w.dim() == 2,
                "w must be a (kh, kw) float32 device tensor");
    w = w.contiguous();
    auto y = torch::empty_like(x, torch::MemoryFormat::Contiguous);
    hipacc_local_desc d;
    std::memset(&d, 0, sizeof(d));
    d.in = view_of(x);
    d.out = view_of(y);
    d.mask_w = (int32_t)w.size(1);
    d.mask_h = (int32_t)w.size(0);
    d.weights_dev = w.data_ptr<float>();
    d.use_weights = 1;
    d.acc_dtype = HIPACC_F32;
    d.reduce = HIPACC_REDUCE_SUM;
    d.boundary = (int32_t)boundary;
    d.const_val = const_val;
    d.cvt = HIPACC_CVT_TRUNC;
    d.fp_contract = 1;
    check(hipacc_hip_local(&d, current_stream()), "hipacc_hip_local");
    ctx->save_for_backward({x, w});
    ctx->saved_data["boundary"] = boundary;
    ctx->saved_data["const_val"] = const_val;
    return y;
  }

  static tensor_list backward(AutogradContext* ctx, tensor_list grads) {
    auto saved = ctx->get_saved_variables();
    auto x = saved[0], w = saved[1];
    auto gy = grads[0].contiguous();
    torch::Tensor gx, gw;
    hipacc_local_grad_desc d;
    std::memset(&d, 0, sizeof(d));
    d.in = view_of(x);
    d.grad_out = view_of(gy);
    if (ctx->needs_input_grad(0)) {
      gx = torch::empty_like(x, torch::MemoryFormat::Contiguous);
      d.grad_in = view_of(gx);
    }
    if (ctx->needs_input_grad(1)) {
      gw = torch::empty_like(w);
      d.grad_weights_dev = gw.data_ptr<float>();
    }
    d.mask_w = (int32_t)w.size(1);
    d.mask_h = (int32_t)w.size(0);
    d.weights_dev = w.data_ptr<float>();
    d.boundary = (int32_t)ctx->saved_data["boundary"].toInt();
    d.const_val = ctx->saved_data["const_val"].toDouble();
    check(hipacc_hip_local_grad(&d, current_stream()), "hipacc_hip_local_grad");
    return {gx, gw, torch::Tensor(), torch::Tensor()};
  }
};

class LearnableBilateral : public torch::autograd::Function<LearnableBilateral> {
 public:
  static torch::Tensor forward(AutogradContext* ctx, torch::Tensor x, torch::Tensor sigma_r,
                               int64_t ksize, double sigma_d, int64_t boundary,
                               double const_val) {
    check_image(x, "x");
    TORCH_CHECK(sigma_r.numel() == 1, "sigma_r must hold one value");
    const float sr = sigma_r.detach().to(torch::kCPU, torch::kFloat32).item<float>();
    auto y = torch::empty_like(x, torch::MemoryFormat::Contiguous);
    auto norm = torch::empty_like(y);
    hipacc_bilateral_grad_desc d;
    std::memset(&d, 0, sizeof(d));
    d.in = view_of(x);
    d.out = view_of(y);
    d.norm = view_of(norm);
    d.ksize = (int32_t)ksize;
    d.sigma_d = (float)sigma_d;
    d.sigma_r = sr;
    d.boundary = (int32_t)boundary;
    d.const_val = const_val;
    check(hipacc_hip_bilateral_train(&d, current_stream()), "hipacc_hip_bilateral_train");
    ctx->save_for_backward({x, y, norm, sigma_r});
    ctx->saved_data["sigma_r"] = (double)sr;
    ctx->saved_data["ksize"] = ksize;
    ctx->saved_data["sigma_d"] = sigma_d;
    ctx->saved_data["boundary"] = boundary;
    ctx->saved_data["const_val"] = const_val;
    return y;
  }

  static tensor_list backward(AutogradContext* ctx, tensor_list grads) {
    auto saved = ctx->get_saved_variables();
    auto x = saved[0], y = saved[1], norm = saved[2], sigma_r = saved[3];
    auto gy = grads[0].contiguous();
    torch::Tensor gx, gs;
    hipacc_bilateral_grad_desc d;
    std::memset(&d, 0, sizeof(d));
    d.in = view_of(x);
    d.out = view_of(y);
    d.norm = view_of(norm);
    d.grad_out = view_of(gy);
    if (ctx->needs_input_grad(0)) {
      gx = torch::empty_like(x, torch::MemoryFormat::Contiguous);
      d.grad_in = view_of(gx);
    }
    if (ctx->needs_input_grad(1)) {
      gs = torch::empty({}, x.options());
      d.grad_sigma_r_dev = gs.data_ptr<float>();
    }
    d.ksize = (int32_t)ctx->saved_data["ksize"].toInt();
    d.sigma_d = (float)ctx->saved_data["sigma_d"].toDouble();
    d.sigma_r = (float)ctx->saved_data["sigma_r"].toDouble();
    d.boundary = (int32_t)ctx->saved_data["boundary"].toInt();
    d.const_val = ctx->saved_data["const_val"].toDouble();
    check(hipacc_hip_bilateral_grad(&d, current_stream()), "hipacc_hip_bilateral_grad");
    if (gs.defined()) gs = gs.to(sigma_r.device(), sigma_r.scalar_type()).reshape(sigma_r.sizes());
    return {gx, gs, torch::Tensor(), torch::Tensor(), torch::Tensor(), torch::Tensor()};
  }
};

}  // namespace

PYBIND11_MODULE(TORCH_EXTENSION_NAME, m) {
  m.doc() = "autograd Functions over the hipacc gfx950 kernels";
  m.def("conv2d",
        [](torch::Tensor x, torch::Tensor w, int64_t boundary, double const_val) {
          return LearnableConv::apply(x, w, boundary, const_val);
        },
        "learnable float convolution (weighted SUM local operator)", py::arg("x"), py::arg("w"),
        py::arg("boundary"), py::arg("const_val") = 0.0);
  m.def("bilateral",
        [](torch::Tensor x, torch::Tensor sigma_r, int64_t ksize, double sigma_d,
           int64_t boundary, double const_val) {
          return LearnableBilateral::apply(x, sigma_r, ksize, sigma_d, boundary, const_val);
        },
        "bilateral filter with learnable range sigma", py::arg("x"), py::arg("sigma_r"),
        py::arg("ksize"), py::arg("sigma_d"), py::arg("boundary"), py::arg("const_val") = 0.0);
}
