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


// The ring of extended positions around a w x h image that boundary
// handling folds back onto it: rows [-ry, 0) and [h, h+ry) over columns
// [-rx, w+rx), and columns [-rx, 0) and [w, w+rx) over rows [0, h).
HH_HD long long ring_size(int w, int h, int rx, int ry) {
  return 2LL * ry * (w + 2 * rx) + 2LL * h * rx;
}
HH_HD long long ring_index(int w, int h, int rx, int ry, int ix, int iy) {
  const long long ew = w + 2 * rx;
  if (iy < 0) return (long long)(iy + ry) * ew + (ix + rx);
  if (iy >= h) return ry * ew + (long long)(iy - h) * ew + (ix + rx);
  const long long base = 2LL * ry * ew;
  if (ix < 0) return base + (long long)iy * rx + (ix + rx);
  return base + (long long)h * rx + (long long)iy * rx + (ix - w);
}
// I: the index type (32-bit where the ring of one image has fewer than
// 2^31 positions).
template <typename I>
HH_HD void ring_pos(int w, int h, int rx, int ry, I j, int& ix, int& iy) {
  const I ew = w + 2 * rx;
  if (j < (I)2 * ry * ew) {
    const I r = j / ew;
    ix = (int)(j - r * ew) - rx;
    iy = r < ry ? (int)r - ry : h + (int)(r - ry);
    return;
  }
  j -= (I)2 * ry * ew;
  const I side = (I)h * rx;
  const bool right = j >= side;
  if (right) j -= side;
  iy = (int)(j / rx);
  const int c = (int)(j - (I)iy * rx);
  ix = right ? w + c : c - rx;
}

// Pixels within (rx, ry) of an edge: full rows at the top and bottom, the
// side columns in between.
HH_HD void band_rows(int w, int h, int rx, int ry, int& top, int& bot0, int& side) {
  top = ry < h ? ry : h;
  bot0 = h - ry > top ? h - ry : top;
  side = rx < w ? rx : w;
}
HH_HD long long band_size(int w, int h, int rx, int ry) {
  int top, bot0, side;
  band_rows(w, h, rx, ry, top, bot0, side);
  const int lcols = side, rcols = w - (w - rx > side ? w - rx : side);
  return (long long)(top + (h - bot0)) * w + (long long)(bot0 - top) * (lcols + rcols);
}
// I: the index type, 32-bit where the band of one image has fewer than 2^31
// pixels (the divisions then take a few instructions instead of a 64-bit
// division sequence).
template <typename I>
HH_HD void band_pos(int w, int h, int rx, int ry, I j, int& x, int& y) {
  int top, bot0, side;
  band_rows(w, h, rx, ry, top, bot0, side);
  const I full = (I)(top + (h - bot0)) * w;
  if (j < full) {
    const int r = (int)(j / w);
    x = (int)(j - (I)r * w);
    y = r < top ? r : bot0 + (r - top);
    return;
  }
  j -= full;
  const int rstart = w - rx > side ? w - rx : side;
  const int lcols = side, rcols = w - rstart, per = lcols + rcols;
  y = top + (int)(j / per);
  const int c = (int)(j - (I)(y - top) * per);
  x = c < lcols ? c : rstart + (c - lcols);
}

// out[t] = scale * SUM_b partials[b * kk + t], t < kk, summed in a fixed
// order (deterministic run to run).
struct PartialSum {
  const float* partials;
  int nblocks, kk;
  float scale;
  float* out;
};
// Block t of a reduction launch (256 threads) computes out[t].
__device__ __forceinline__ void partial_sum_block(const PartialSum& p, int t) {
  __shared__ float red[4];
  float s = 0.0f;
  for (int b = threadIdx.x; b < p.nblocks; b += 256) s += p.partials[(long long)b * p.kk + t];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) p.out[t] = (((red[0] + red[1]) + red[2]) + red[3]) * p.scale;
}
int launch_partial_sum(const PartialSum& p, hipStream_t s);

// Adds, for every band pixel q, the ring values of the extended positions
// (other than q) its boundary mode maps onto q, in a fixed order.  With
// `sum` (not NULL), the same launch also performs that reduction (the
// partials are complete when the fold starts).
int launch_fold_add(float* gx, long long gx_plane, int gx_stride, int w, int h, int batch, int rx,
                    int ry, int mode, const float* ring, hipStream_t s,
                    const PartialSum* sum = nullptr);

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
  int dw_variant;      // kVarTile: LDS tiles; otherwise row streaming
};

// E(i) = SUM_t w[t] * dY0(i + r - t) (dY0: dY inside the image, 0 outside)
// at every ring position i, into ring[b * ring_size + ring_index(i)].
// With wflip (not NULL) the same launch writes the flipped device mask
// wflip[t] = wdev[kx*ky-1-t] for the dX pass.
int launch_conv_ring(const ConvGradArgs& a, float* ring, float* wflip, hipStream_t s);
// Per-block partial sums of dw into a.partials (the reduction, `sum`, is
// left to the caller: launch_partial_sum or a fold launch).
int launch_conv_dw(const ConvGradArgs& a, float* dw, hipStream_t s, PartialSum* sum);
// out[t] = w[n-1-t] (the flipped mask of the dX correlation).
int launch_flip(const float* w, float* out, int n, hipStream_t s);

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
  float lsp[kMaxTaps]; // log2 of the spatial weights
  int sym;             // sp is point-symmetric
  float* partials;     // [blocks] partial sums of dL/dc
  float* ring;         // neighbour terms of the ring positions (NULL: none)
  long long rsize;     // ring positions per image
  int vec_ok;
  int tiles_x, tiles_y, ntiles;
};
// dsigma (device, 1 float) = -1/sigma_r^3 * dL/dc when not NULL; ring
// (when not NULL) receives the folded neighbour terms of the ring positions
// and is added onto the band pixels of dX.
int launch_bilateral_grad(const BilatGradArgs& a, float* dsigma, float sigma_r, float* ring,
                          hipStream_t s);

}  // namespace hh
