// hh_grad.hpp -- argument blocks of the adjoint (backward) kernels of the
// float weighted-SUM local operator and the bilateral filter.
#pragma once

#include "hh_local.hpp"

namespace hh {

constexpr int kGradBlocks = 2048;   // max blocks writing weight-gradient partials

// Positions i in [-r, n-1+r] the boundary map sends to q; q itself first.
// The image is the whole allocation, so UNDEFINED behaves as CLAMP.
HH_HD int preimage(int mode, int q, int n, int r, int* out) {
  int k = 0;
  out[k++] = q;
  if (mode == HIPACC_BOUNDARY_CONSTANT) return k;
  for (int i = -r; i < 0; ++i)
    if (bh_map(mode, i, 0, n, n) == q) out[k++] = i;
  for (int i = n; i < n + r; ++i)
    if (bh_map(mode, i, 0, n, n) == q) out[k++] = i;
  return k;
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}


// Adjoint of out(p) = SUM_t w[t] * X_bh(p + t - r) over whole images (the
// accessor window is the image, the iteration space has its size).
struct ConvGradArgs {
  const float* x;      // forward input X
  const float* gy;     // dL/dY
  float* gx;           // dL/dX (NULL: not computed)
  long long x_plane, gy_plane, gx_plane;
  int x_stride, gy_stride, gx_stride;
  int w, h, batch;
  int kx, ky;
  int boundary;
  float cval;          // CONSTANT value (contributes to dL/dw)
  int vec_ok;          // every row of x, gy, gx is 16-byte aligned
  const float* wdev;   // device weights, or NULL: wh
  float wh[kMaxTaps];
  float* partials;     // [blocks][kx*ky] weight-gradient partial sums
  int tiles_x, tiles_y, ntiles;
};

// Launches the dX / dW kernels and, when dw != NULL, the final reduction of
// the partials into dw (device, kx*ky floats).
int launch_conv_grad(const ConvGradArgs& a, float* dw, hipStream_t s);

// Backward of the bilateral filter (hipacc_hip_bilateral_grad).
struct BilatGradArgs {
  const float* x;      // forward input
  const float* y;      // forward output
  const float* d;      // forward normaliser SUM_t s_t
  const float* gy;     // dL/dY
  float* gx;           // dL/dX (NULL: not computed)
  long long x_plane, y_plane, d_plane, gy_plane, gx_plane;
  int x_stride, y_stride, d_stride, gy_stride, gx_stride;
  int w, h, batch;
  int ksize;
  int boundary;
  float cval;
  float c;             // 0.5 / sigma_r^2
  float k2;            // -c * log2(e)
  float sp[kMaxTaps];  // spatial weights
  float* partials;     // [blocks] partial sums of dL/dc
  int vec_ok;
  int tiles_x, tiles_y, ntiles;
};
// dsigma (device, 1 float) = -1/sigma_r^3 * dL/dc when not NULL.
int launch_bilateral_grad(const BilatGradArgs& a, float* dsigma, float sigma_r, hipStream_t s);

}  // namespace hh
