// grad_conv.hip -- adjoint of the float weighted-SUM local operator
//     Y(p) = SUM_t w[t] * X_bh(p + t - r)           (t row-major, r = mask radius)
// over whole images with boundary mode bh:
//     dX(q) = SUM_{p,t : bh(p + t - r) = q} w[t] * dY(p)
//     dw[t] = SUM_p dY(p) * X_bh(p + t - r)
// dX is computed in gather form.  With E(i) = SUM_t w[t] * dY0(i - t + r)
// (dY0 = dY inside the image, 0 outside) over the extended index range
// i in [-r, n-1+r], and P(q) = { i : bh(i) = q } (separable per axis),
//     dX(q) = SUM_{i in P(q)} E(i).
// Away from the borders P(q) = {q} and dX is the flipped-mask correlation of
// dY; pixels within r of an edge add the folded terms of the halo positions
// that the boundary mode maps onto them (CLAMP: the run of clamped positions,
// MIRROR: the reflected position, REPEAT: the wrapped one; CONSTANT: none).
// dw is a per-tap reduction: per-block partial sums, then a fixed-order
// reduction (deterministic run to run).
#include "hh_grad.hpp"

namespace hh {

__device__ __forceinline__ float grad_weight(const ConvGradArgs& a, int t) {
  return a.wdev ? a.wdev[t] : a.wh[t];
}

// E(ix, iy) read from global memory (any extended position).
__device__ float conv_e(const ConvGradArgs& a, const float* G, int ix, int iy) {
  const int rx = a.kx / 2, ry = a.ky / 2;
  float e = 0.0f;
  for (int ty = 0; ty < a.ky; ++ty) {
    const int y = iy + ry - ty;
    if (y < 0 || y >= a.h) continue;
    for (int tx = 0; tx < a.kx; ++tx) {
      const int x = ix + rx - tx;
      if (x < 0 || x >= a.w) continue;
      e = __builtin_fmaf(grad_weight(a, ty * a.kx + tx), G[(long long)y * a.gy_stride + x], e);
    }
  }
  return e;
}

// SUM of E over P(qx) x P(qy), without (qx, qy) itself when skip_self.
__device__ float conv_fold(const ConvGradArgs& a, const float* G, int qx, int qy, bool skip_self) {
  int px[2 * (HIPACC_MAX_MASK_DIM / 2) + 1], py[2 * (HIPACC_MAX_MASK_DIM / 2) + 1];
  const int nx = preimage(a.boundary, qx, a.w, a.kx / 2, px);
  const int ny = preimage(a.boundary, qy, a.h, a.ky / 2, py);
  float acc = 0.0f;
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      if (skip_self && i == 0 && j == 0) continue;
      acc += conv_e(a, G, px[i], py[j]);
    }
  return acc;
}

// dX and/or dw over 64 x 16 pixel tiles of square K = 2R+1 masks.  Each
// lane owns 4 adjacent pixels of one row; dY and X tiles with halo are
// staged in LDS; a block walks tiles blockIdx.x, +gridDim.x, ... and keeps
// its dw partial sums (K*K per lane) in registers.
template <int R, bool DX, bool DW>
__global__ __launch_bounds__(256) void conv_grad_kernel(const ConvGradArgs a) {
  constexpr int K = 2 * R + 1, KK = K * K;
  constexpr int TW = 64, TH = 16;
  constexpr int RH = (R + 3) & ~3;   // column halo rounded to 16-byte groups
  constexpr int LW = TW + 2 * RH, LH = TH + 2 * R, LW4 = LW / 4;
  constexpr int M = 4 + 2 * R;
  __shared__ __align__(16) float gs[LH * LW];
  __shared__ __align__(16) float xs[DW ? LH * LW : 4];
  __shared__ float red[DW ? 4 * KK : 4];

  const int tid = threadIdx.x, tr = tid >> 4, tc = (tid & 15) * 4;
  float dw[DW ? KK : 1];
#pragma unroll
  for (int t = 0; t < (DW ? KK : 1); ++t) dw[t] = 0.0f;
  const int per_plane = a.tiles_x * a.tiles_y;

  for (int tile = blockIdx.x; tile < a.ntiles; tile += gridDim.x) {
    const int b = tile / per_plane, rem = tile - b * per_plane;
    const int y0 = (rem / a.tiles_x) * TH, x0 = (rem % a.tiles_x) * TW;
    const float* X = a.x + b * a.x_plane;
    const float* G = a.gy + b * a.gy_plane;
    const bool interior =
        a.vec_ok && x0 >= RH && x0 + TW + RH <= a.w && y0 >= R && y0 + TH + R <= a.h;
    __syncthreads();  // the previous tile's LDS reads are done
    if (interior) {
      constexpr int N4 = LH * LW4;
      for (int i = tid; i < N4; i += 256) {
        const int r = i / LW4, c = (i - r * LW4) * 4;
        const long long col = x0 - RH + c;
        const int row = y0 - R + r;
        const float4 gv = *reinterpret_cast<const float4*>(G + (long long)row * a.gy_stride + col);
        if constexpr (DW) {
          const float4 xv = *reinterpret_cast<const float4*>(X + (long long)row * a.x_stride + col);
          *reinterpret_cast<float4*>(xs + r * LW + c) = xv;
        }
        *reinterpret_cast<float4*>(gs + r * LW + c) = gv;
      }
    } else {
      for (int i = tid; i < LH * LW; i += 256) {
        const int r = i / LW, c = i - r * LW;
        const int y = y0 - R + r, x = x0 - RH + c;
        const bool inside = x >= 0 && x < a.w && y >= 0 && y < a.h;
        gs[i] = inside ? G[(long long)y * a.gy_stride + x] : 0.0f;
        if constexpr (DW) {
          float v;
          if (inside)
            v = X[(long long)y * a.x_stride + x];
          else if (a.boundary == HIPACC_BOUNDARY_CONSTANT)
            v = a.cval;
          else
            v = X[(long long)bh_map(a.boundary, y, 0, a.h, a.h) * a.x_stride +
                  bh_map(a.boundary, x, 0, a.w, a.w)];
          xs[i] = v;
        }
      }
    }
    __syncthreads();

    float gx[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    float gc[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) gc[j] = gs[(tr + R) * LW + tc + RH + j];
#pragma unroll
    for (int dy = 0; dy < K; ++dy) {
      float gr[M], xr[DW ? M : 1];
      const int off = (tr + dy) * LW + tc + (RH - R);
#pragma unroll
      for (int m = 0; m < M; ++m) {
        gr[m] = gs[off + m];
        if constexpr (DW) xr[DW ? m : 0] = xs[off + m];
      }
#pragma unroll
      for (int dx = 0; dx < K; ++dx) {
        if constexpr (DX) {
          const float wv = grad_weight(a, (K - 1 - dy) * K + (K - 1 - dx));
#pragma unroll
          for (int j = 0; j < 4; ++j) gx[j] = __builtin_fmaf(wv, gr[j + dx], gx[j]);
        }
        if constexpr (DW) {
#pragma unroll
          for (int j = 0; j < 4; ++j)
            dw[DW ? dy * K + dx : 0] =
                __builtin_fmaf(gc[j], xr[DW ? j + dx : 0], dw[DW ? dy * K + dx : 0]);
        }
      }
    }

    if constexpr (DX) {
      const int qy = y0 + tr;
      if (qy < a.h) {
        float* GX = a.gx + b * a.gx_plane + (long long)qy * a.gx_stride;
        if (interior) {
          *reinterpret_cast<float4*>(GX + x0 + tc) = make_float4(gx[0], gx[1], gx[2], gx[3]);
        } else {
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const int qx = x0 + tc + j;
            if (qx >= a.w) continue;
            float v = gx[j];
            if (a.boundary != HIPACC_BOUNDARY_CONSTANT &&
                (qx < R || qx >= a.w - R || qy < R || qy >= a.h - R))
              v += conv_fold(a, G, qx, qy, true);
            GX[qx] = v;
          }
        }
      }
    }
  }

  if constexpr (DW) {
    const int lane = tid & 63, wave = tid >> 6;
#pragma unroll
    for (int t = 0; t < (DW ? KK : 1); ++t) {
      const float v = wave_sum(dw[t]);
      if (lane == 0) red[wave * KK + t] = v;
    }
    __syncthreads();
    for (int t = tid; t < KK; t += 256)
      a.partials[(long long)blockIdx.x * KK + t] =
          ((red[t] + red[KK + t]) + red[2 * KK + t]) + red[3 * KK + t];
  }
}

// Any mask shape: one thread per dX pixel, full preimage sum.
__global__ __launch_bounds__(256) void conv_grad_dx_generic(const ConvGradArgs a) {
  const int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y, b = blockIdx.z;
  if (x >= a.w) return;
  const float* G = a.gy + b * a.gy_plane;
  a.gx[b * a.gx_plane + (long long)y * a.gx_stride + x] = conv_fold(a, G, x, y, false);
}

// Any mask shape: block (chunk, tap) sums dY(p) * X_bh(p + t - r) over the
// pixels p = chunk*256 + tid + k*nchunks*256 (all images).
__global__ __launch_bounds__(256) void conv_grad_dw_generic(const ConvGradArgs a) {
  __shared__ float red[4];
  const int kk = a.kx * a.ky, t = blockIdx.y;
  const int ty = t / a.kx, tx = t - ty * a.kx, rx = a.kx / 2, ry = a.ky / 2;
  const long long npx = (long long)a.w * a.h, total = npx * a.batch;
  float acc = 0.0f;
  for (long long p = (long long)blockIdx.x * 256 + threadIdx.x; p < total;
       p += (long long)gridDim.x * 256) {
    const int b = (int)(p / npx);
    const long long r = p - b * npx;
    const int y = (int)(r / a.w), x = (int)(r - (long long)y * a.w);
    const float g = a.gy[b * a.gy_plane + (long long)y * a.gy_stride + x];
    int sx = x + tx - rx, sy = y + ty - ry;
    float v;
    if (sx >= 0 && sx < a.w && sy >= 0 && sy < a.h) {
      v = a.x[b * a.x_plane + (long long)sy * a.x_stride + sx];
    } else if (a.boundary == HIPACC_BOUNDARY_CONSTANT) {
      v = a.cval;
    } else {
      sx = bh_map(a.boundary, sx, 0, a.w, a.w);
      sy = bh_map(a.boundary, sy, 0, a.h, a.h);
      v = a.x[b * a.x_plane + (long long)sy * a.x_stride + sx];
    }
    acc = __builtin_fmaf(g, v, acc);
  }
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0)
    a.partials[(long long)blockIdx.x * kk + t] = ((red[0] + red[1]) + red[2]) + red[3];
}

// dw[t] = SUM_b partials[b][t] in a fixed order.
__global__ __launch_bounds__(256) void conv_grad_dw_reduce(const float* partials, int nblocks,
                                                           int kk, float* dw) {
  __shared__ float red[4];
  const int t = blockIdx.x;
  float s = 0.0f;
  for (int b = threadIdx.x; b < nblocks; b += 256) s += partials[(long long)b * kk + t];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) dw[t] = ((red[0] + red[1]) + red[2]) + red[3];
}

template <int R>
static int launch_conv_grad_fast(const ConvGradArgs& a, bool dx, bool dw, int nblocks,
                                 hipStream_t s) {
  // dX and dw run as two launches: fused, the K*K dw accumulators and the dX
  // windows together exceed the register file (occupancy 1); apart each
  // keeps 4 waves per SIMD, and the second pass re-reads dY from cache.
  if (dx) {
    hipLaunchKernelGGL((conv_grad_kernel<R, true, false>), dim3(nblocks), dim3(256), 0, s, a);
    int st = hip_status(hipGetLastError(), "conv_grad_kernel launch");
    if (st) return st;
  }
  if (dw) {
    hipLaunchKernelGGL((conv_grad_kernel<R, false, true>), dim3(nblocks), dim3(256), 0, s, a);
    return hip_status(hipGetLastError(), "conv_grad_kernel launch");
  }
  return HIPACC_OK;
}

int launch_conv_grad(const ConvGradArgs& a0, float* dw, hipStream_t s) {
  ConvGradArgs a = a0;
  const bool dx = a.gx != nullptr, dwn = dw != nullptr;
  if (!dx && !dwn) return HIPACC_OK;
  const int kk = a.kx * a.ky;
  int nblocks = 0;
  int st = HIPACC_OK;
  if (a.kx == a.ky && (a.kx == 3 || a.kx == 5 || a.kx == 7 || a.kx == 9)) {
    a.tiles_x = ceil_div(a.w, 64);
    a.tiles_y = ceil_div(a.h, 16);
    a.ntiles = a.tiles_x * a.tiles_y * a.batch;
    nblocks = a.ntiles < kGradBlocks ? a.ntiles : kGradBlocks;
    switch (a.kx) {
      case 3: st = launch_conv_grad_fast<1>(a, dx, dwn, nblocks, s); break;
      case 5: st = launch_conv_grad_fast<2>(a, dx, dwn, nblocks, s); break;
      case 7: st = launch_conv_grad_fast<3>(a, dx, dwn, nblocks, s); break;
      default: st = launch_conv_grad_fast<4>(a, dx, dwn, nblocks, s); break;
    }
  } else {
    if (dx) {
      hipLaunchKernelGGL(conv_grad_dx_generic, dim3(ceil_div(a.w, 256), a.h, a.batch), dim3(256),
                         0, s, a);
      st = hip_status(hipGetLastError(), "conv_grad_dx_generic launch");
    }
    if (st == HIPACC_OK && dwn) {
      const long long total = (long long)a.w * a.h * a.batch;
      long long nc = (total + 256 * 16 - 1) / (256 * 16);
      const int cap = 256;
      nblocks = (int)(nc < cap ? nc : cap);
      hipLaunchKernelGGL(conv_grad_dw_generic, dim3(nblocks, kk), dim3(256), 0, s, a);
      st = hip_status(hipGetLastError(), "conv_grad_dw_generic launch");
    }
  }
  if (st == HIPACC_OK && dwn) {
    hipLaunchKernelGGL(conv_grad_dw_reduce, dim3(kk), dim3(256), 0, s, a.partials, nblocks, kk,
                       dw);
    st = hip_status(hipGetLastError(), "conv_grad_dw_reduce launch");
  }
  return st;
}

}  // namespace hh
