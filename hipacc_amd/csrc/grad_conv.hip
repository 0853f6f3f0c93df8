// grad_conv.hip -- adjoint of the float weighted-SUM local operator
//     Y(p) = SUM_t w[t] * X_bh(p + t - r)           (t row-major, r = mask radius)
// over whole images with boundary mode bh:
//     dX(q) = SUM_{p,t : bh(p + t - r) = q} w[t] * dY(p)
//     dw[t] = SUM_p dY(p) * X_bh(p + t - r)
// dX in gather form: with E(i) = SUM_t w[t] * dY0(i + r - t) (dY0 = dY inside
// the image, 0 outside) over the extended index range, and P(q) = { i :
// bh(i) = q } (separable per axis),
//     dX(q) = SUM_{i in P(q)} E(i).
// E over the image is the forward operator itself applied to dY with the
// flipped mask and CONSTANT 0 borders (the caller runs it through the
// forward's kernels).  The rest of P(q) -- the halo positions the mode folds
// onto pixels within r of an edge (CLAMP: the clamped run, MIRROR: the
// reflection, REPEAT: the wrap; CONSTANT: none) -- is evaluated once per
// ring position here and added to the border band.  dw: per-block partial
// sums, then a fixed-order reduction (deterministic run to run).
#include <algorithm>

#include "hh_grad.hpp"
#include "hh_stream.hpp"

namespace hh {

__device__ __forceinline__ float grad_weight(const ConvGradArgs& a, int t) {
  return a.wdev ? ((const HH_CONST float*)a.wdev)[t] : a.wh[t];
}

__global__ void flip_kernel(const float* w, float* out, int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) out[t] = w[n - 1 - t];
}

int launch_flip(const float* w, float* out, int n, hipStream_t s) {
  hipLaunchKernelGGL(flip_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, w, out, n);
  return hip_status(hipGetLastError(), "flip_kernel launch");
}

// One thread per ring position of one image (KX, KY: the mask size when it
// is a compile-time constant, unrolled with the taps outside the image
// masked; 0: any size).  Block (0, 0) also writes the flipped mask.
template <int KX, int KY, typename I>
__global__ __launch_bounds__(256) void conv_ring_kernel(const ConvGradArgs a, float* ring,
                                                        long long rsize, float* wflip) {
  const int kx = KX ? KX : a.kx, ky = KY ? KY : a.ky;
  const int b = blockIdx.y;
  if (wflip && b == 0 && blockIdx.x == 0)
    for (int t = threadIdx.x; t < kx * ky; t += 256) wflip[t] = grad_weight(a, kx * ky - 1 - t);
  const I j = (I)blockIdx.x * 256 + threadIdx.x;
  if (j >= (I)rsize) return;
  const int rx = kx / 2, ry = ky / 2;
  int ix, iy;
  ring_pos<I>(a.w, a.h, rx, ry, j, ix, iy);
  const float* G = a.gy + b * a.gy_plane;
  float e = 0.0f;
  if constexpr (KX > 0 && KY > 0) {
    // the loads of a tap row issued together (rows outside the image, the
    // same for the lanes of a wave in the top and bottom bands, skipped);
    // columns clamped into the image, the taps outside it leave e unchanged
#pragma unroll
    for (int ty = 0; ty < KY; ++ty) {
      const int y = iy + ry - ty;
      if (y < 0 || y >= a.h) continue;
      const float* row = G + (long long)y * a.gy_stride;
      float v[KX];
#pragma unroll
      for (int tx = 0; tx < KX; ++tx) v[tx] = row[min(max(ix + rx - tx, 0), a.w - 1)];
#pragma unroll
      for (int tx = 0; tx < KX; ++tx) {
        const int x = ix + rx - tx;
        const float f = __builtin_fmaf(grad_weight(a, ty * KX + tx), v[tx], e);
        e = (x >= 0 && x < a.w) ? f : e;
      }
    }
  } else {
    for (int ty = 0; ty < ky; ++ty) {
      const int y = iy + ry - ty;
      if (y < 0 || y >= a.h) continue;
      for (int tx = 0; tx < kx; ++tx) {
        const int x = ix + rx - tx;
        if (x < 0 || x >= a.w) continue;
        e = __builtin_fmaf(grad_weight(a, ty * kx + tx), G[(long long)y * a.gy_stride + x], e);
      }
    }
  }
  ring[b * rsize + (long long)j] = e;
}

int launch_conv_ring(const ConvGradArgs& a, float* ring, float* wflip, hipStream_t s) {
  const long long rs = ring_size(a.w, a.h, a.kx / 2, a.ky / 2);
  if (rs == 0) return wflip ? launch_flip(a.wdev, wflip, a.kx * a.ky, s) : HIPACC_OK;
  dim3 grid((unsigned)((rs + 255) / 256), a.batch);
  const bool small = rs + 256 < (1LL << 31);
  if (small && a.kx == a.ky && (a.kx == 3 || a.kx == 5 || a.kx == 7 || a.kx == 9)) {
    switch (a.kx) {
      case 3: hipLaunchKernelGGL((conv_ring_kernel<3, 3, int>), grid, dim3(256), 0, s, a, ring, rs, wflip); break;
      case 5: hipLaunchKernelGGL((conv_ring_kernel<5, 5, int>), grid, dim3(256), 0, s, a, ring, rs, wflip); break;
      case 7: hipLaunchKernelGGL((conv_ring_kernel<7, 7, int>), grid, dim3(256), 0, s, a, ring, rs, wflip); break;
      default: hipLaunchKernelGGL((conv_ring_kernel<9, 9, int>), grid, dim3(256), 0, s, a, ring, rs, wflip); break;
    }
  } else if (small) {
    hipLaunchKernelGGL((conv_ring_kernel<0, 0, int>), grid, dim3(256), 0, s, a, ring, rs, wflip);
  } else {
    hipLaunchKernelGGL((conv_ring_kernel<0, 0, long long>), grid, dim3(256), 0, s, a, ring, rs,
                       wflip);
  }
  return hip_status(hipGetLastError(), "conv_ring_kernel launch");
}

// One thread per band pixel of one image; with a reduction, kk more blocks
// (of row 0 of the grid) sum the partials.  RX, RY > 0: the mask radii as
// compile-time constants: the candidate positions of each axis (q, then
// [-r, 0), then [n, n + r)) are unrolled and every ring load is issued
// before the sum, in the order of the general loop below.
template <int RX, int RY>
__device__ __forceinline__ float fold_sum_fixed(int mode, int qx, int qy, int w, int h,
                                                const float* R, bool& any) {
  constexpr int NX = 2 * RX + 1, NY = 2 * RY + 1;
  int cx[NX], cy[NY];
  bool mx[NX], my[NY];
  cx[0] = qx;
  mx[0] = true;
#pragma unroll
  for (int k = 0; k < RX; ++k) {
    cx[1 + k] = k - RX;
    cx[1 + RX + k] = w + k;
  }
#pragma unroll
  for (int k = 1; k < NX; ++k) mx[k] = bh_map(mode, cx[k], 0, w, w) == qx;
  cy[0] = qy;
  my[0] = true;
#pragma unroll
  for (int k = 0; k < RY; ++k) {
    cy[1 + k] = k - RY;
    cy[1 + RY + k] = h + k;
  }
#pragma unroll
  for (int k = 1; k < NY; ++k) my[k] = bh_map(mode, cy[k], 0, h, h) == qy;
  float v[NY][NX];
  any = false;
#pragma unroll
  for (int jy = 0; jy < NY; ++jy)
#pragma unroll
    for (int jx = 0; jx < NX; ++jx) {
      const bool on = my[jy] && mx[jx] && (jx | jy) != 0;
      any |= on;
      v[jy][jx] = on ? R[ring_index(w, h, RX, RY, cx[jx], cy[jy])] : 0.0f;
    }
  float acc = 0.0f;
#pragma unroll
  for (int jy = 0; jy < NY; ++jy)
#pragma unroll
    for (int jx = 0; jx < NX; ++jx)
      if (my[jy] && mx[jx] && (jx | jy) != 0) acc += v[jy][jx];
  return acc;
}

template <typename I, int RX, int RY>
__global__ __launch_bounds__(256) void fold_add_kernel(float* gx, long long gx_plane,
                                                       int gx_stride, int w, int h, int rx,
                                                       int ry, int mode, const float* ring,
                                                       long long rsize, long long bsize,
                                                       int fold_blocks, const PartialSum sum) {
  const int b = blockIdx.y;
  if ((int)blockIdx.x >= fold_blocks) {
    if (b == 0) partial_sum_block(sum, (int)blockIdx.x - fold_blocks);
    return;
  }
  const I j = (I)blockIdx.x * 256 + threadIdx.x;
  if (j >= (I)bsize) return;
  int qx, qy;
  band_pos<I>(w, h, rx, ry, j, qx, qy);
  const float* R = ring + b * rsize;
  if constexpr (RX > 0 && RY > 0) {
    bool any;
    const float acc = fold_sum_fixed<RX, RY>(mode, qx, qy, w, h, R, any);
    if (!any) return;
    float* dst = gx + b * gx_plane + (long long)qy * gx_stride + qx;
    *dst = *dst + acc;
    return;
  }
  int px[2 * (HIPACC_MAX_MASK_DIM / 2) + 1], py[2 * (HIPACC_MAX_MASK_DIM / 2) + 1];
  const int nx = preimage(mode, qx, w, rx, px);
  const int ny = preimage(mode, qy, h, ry, py);
  if (nx == 1 && ny == 1) return;
  float acc = 0.0f;
  for (int jy = 0; jy < ny; ++jy)
    for (int jx = 0; jx < nx; ++jx) {
      if (jx == 0 && jy == 0) continue;
      acc += R[ring_index(w, h, rx, ry, px[jx], py[jy])];
    }
  float* dst = gx + b * gx_plane + (long long)qy * gx_stride + qx;
  *dst = *dst + acc;
}

int launch_fold_add(float* gx, long long gx_plane, int gx_stride, int w, int h, int batch, int rx,
                    int ry, int mode, const float* ring, hipStream_t s, const PartialSum* sum) {
  const long long bs = mode == HIPACC_BOUNDARY_CONSTANT ? 0 : band_size(w, h, rx, ry);
  const long long rs = ring_size(w, h, rx, ry);
  if (bs == 0) return sum ? launch_partial_sum(*sum, s) : HIPACC_OK;
  const int fb = (int)((bs + 255) / 256);
  PartialSum ps{};
  if (sum) ps = *sum;
  dim3 grid((unsigned)(fb + (sum ? sum->kk : 0)), batch);
  const bool small = bs + 256 < (1LL << 31);
  if (small && rx == ry && rx >= 1 && rx <= 4) {
    switch (rx) {
      case 1: hipLaunchKernelGGL((fold_add_kernel<int, 1, 1>), grid, dim3(256), 0, s, gx, gx_plane, gx_stride, w, h, rx, ry, mode, ring, rs, bs, fb, ps); break;
      case 2: hipLaunchKernelGGL((fold_add_kernel<int, 2, 2>), grid, dim3(256), 0, s, gx, gx_plane, gx_stride, w, h, rx, ry, mode, ring, rs, bs, fb, ps); break;
      case 3: hipLaunchKernelGGL((fold_add_kernel<int, 3, 3>), grid, dim3(256), 0, s, gx, gx_plane, gx_stride, w, h, rx, ry, mode, ring, rs, bs, fb, ps); break;
      default: hipLaunchKernelGGL((fold_add_kernel<int, 4, 4>), grid, dim3(256), 0, s, gx, gx_plane, gx_stride, w, h, rx, ry, mode, ring, rs, bs, fb, ps); break;
    }
  } else if (small) {
    hipLaunchKernelGGL((fold_add_kernel<int, 0, 0>), grid, dim3(256), 0, s, gx, gx_plane,
                       gx_stride, w, h, rx, ry, mode, ring, rs, bs, fb, ps);
  } else {
    hipLaunchKernelGGL((fold_add_kernel<long long, 0, 0>), grid, dim3(256), 0, s, gx, gx_plane,
                       gx_stride, w, h, rx, ry, mode, ring, rs, bs, fb, ps);
  }
  return hip_status(hipGetLastError(), "fold_add_kernel launch");
}

__global__ __launch_bounds__(256) void partial_sum_kernel(const PartialSum p) {
  partial_sum_block(p, blockIdx.x);
}

int launch_partial_sum(const PartialSum& p, hipStream_t s) {
  hipLaunchKernelGGL(partial_sum_kernel, dim3(p.kk), dim3(256), 0, s, p);
  return hip_status(hipGetLastError(), "partial_sum_kernel launch");
}

// dw over 64 x 32 pixel tiles of square K = 2R+1 masks.  Each lane owns 4
// adjacent pixels in each of 2 consecutive rows (the two rows share K-1 of
// their K+1 X rows) and keeps K*K partial sums; the X tile with its halo and
// the dY tile are staged in LDS.  A block walks a contiguous range of tiles
// and loads the next interior tile into registers before computing the
// current one.
// Part i of a tile (16-byte group i of LH x LW4 groups) from global memory,
// and into LDS.
template <int LW4, int NX4, int R, int RH>
__device__ __forceinline__ void tile_part_load(float4& p, int i, const float* X, int stride,
                                               int x0, int y0) {
  if (i < NX4) {
    const int r = i / LW4, c = (i - r * LW4) * 4;
    p = *reinterpret_cast<const float4*>(X + (long long)(y0 - R + r) * stride + (x0 - RH + c));
  }
}
template <int LW4, int NX4>
__device__ __forceinline__ void tile_part_store(const float4& p, int i, float* xs) {
  if (i < NX4) {
    const int r = i / LW4, c = (i - r * LW4) * 4;
    *reinterpret_cast<float4*>(xs + r * (LW4 * 4) + c) = p;
  }
}

template <int R>
__global__ __launch_bounds__(256) void conv_dw_kernel(const ConvGradArgs a, int per_block) {
  constexpr int K = 2 * R + 1, KK = K * K;
  constexpr int TW = 64, TH = 32;
  constexpr int RH = (R + 3) & ~3;   // column halo rounded to 16-byte groups
  constexpr int LW = TW + 2 * RH, LH = TH + 2 * R, LW4 = LW / 4;
  constexpr int NX4 = LH * LW4, NL = (NX4 + 255) / 256;
  static_assert(NL <= 3, "tile staging assumes at most 3 parts per thread");
  constexpr int M = 4 + 2 * R;
  __shared__ __align__(16) float xs[LH * LW];
  __shared__ __align__(16) float gs[TH * TW];
  __shared__ float red[4 * KK];

  const int tid = threadIdx.x, tr = (tid >> 4) * 2, tc = (tid & 15) * 4;
  float dw[KK];
#pragma unroll
  for (int t = 0; t < KK; ++t) dw[t] = 0.0f;

  const int t0 = blockIdx.x * per_block;
  const int t1 = min(t0 + per_block, a.ntiles);
  // tile coordinates, advanced incrementally
  int b = 0, ty = 0, tx = 0;
  if (t0 < t1) {
    const int per_plane = a.tiles_x * a.tiles_y;
    b = t0 / per_plane;
    const int rem = t0 - b * per_plane;
    ty = rem / a.tiles_x;
    tx = rem - ty * a.tiles_x;
  }
  auto interior = [&](int x0, int y0) {
    return a.vec_ok && x0 >= RH && x0 + TW + RH <= a.w && y0 >= R && y0 + TH + R <= a.h;
  };
  float4 p0, p1, p2, pg0, pg1;
  auto prefetch = [&](int pb, int x0, int y0) {
    const float* X = a.x + pb * a.x_plane;
    const float* G = a.gy + pb * a.gy_plane;
    tile_part_load<LW4, NX4, R, RH>(p0, tid, X, a.x_stride, x0, y0);
    if constexpr (NL > 1) tile_part_load<LW4, NX4, R, RH>(p1, tid + 256, X, a.x_stride, x0, y0);
    if constexpr (NL > 2) tile_part_load<LW4, NX4, R, RH>(p2, tid + 512, X, a.x_stride, x0, y0);
    pg0 = *reinterpret_cast<const float4*>(G + (long long)(y0 + tr) * a.gy_stride + x0 + tc);
    pg1 = *reinterpret_cast<const float4*>(G + (long long)(y0 + tr + 1) * a.gy_stride + x0 + tc);
  };

  bool cur = t0 < t1 && interior(tx * TW, ty * TH);
  if (cur) prefetch(b, tx * TW, ty * TH);
  for (int t = t0; t < t1; ++t) {
    const int x0 = tx * TW, y0 = ty * TH;
    __syncthreads();  // the previous tile's LDS reads are done
    if (cur) {
      tile_part_store<LW4, NX4>(p0, tid, xs);
      if constexpr (NL > 1) tile_part_store<LW4, NX4>(p1, tid + 256, xs);
      if constexpr (NL > 2) tile_part_store<LW4, NX4>(p2, tid + 512, xs);
      *reinterpret_cast<float4*>(gs + tr * TW + tc) = pg0;
      *reinterpret_cast<float4*>(gs + (tr + 1) * TW + tc) = pg1;
    } else {
      const float* X = a.x + b * a.x_plane;
      const float* G = a.gy + b * a.gy_plane;
      for (int i = tid; i < LH * LW; i += 256) {
        const int r = i / LW, c = i - r * LW;
        const int y = y0 - R + r, x = x0 - RH + c;
        float v;
        if (x >= 0 && x < a.w && y >= 0 && y < a.h)
          v = X[(long long)y * a.x_stride + x];
        else if (a.boundary == HIPACC_BOUNDARY_CONSTANT)
          v = a.cval;
        else
          v = X[(long long)bh_map(a.boundary, y, 0, a.h, a.h) * a.x_stride +
                bh_map(a.boundary, x, 0, a.w, a.w)];
        xs[i] = v;
      }
#pragma unroll
      for (int q = 0; q < 2; ++q)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int y = y0 + tr + q, x = x0 + tc + j;
          gs[(tr + q) * TW + tc + j] =
              (x < a.w && y < a.h) ? G[(long long)y * a.gy_stride + x] : 0.0f;
        }
    }
    __syncthreads();
    // advance to the next tile and prefetch it
    if (++tx == a.tiles_x) {
      tx = 0;
      if (++ty == a.tiles_y) {
        ty = 0;
        ++b;
      }
    }
    const bool nxt = t + 1 < t1 && interior(tx * TW, ty * TH);
    if (nxt) prefetch(b, tx * TW, ty * TH);

    float g0[4], g1[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      g0[j] = gs[tr * TW + tc + j];
      g1[j] = gs[(tr + 1) * TW + tc + j];
    }
#pragma unroll
    for (int rho = 0; rho <= K; ++rho) {
      float xr[M];
      const int off = (tr + rho) * LW + tc + (RH - R);
#pragma unroll
      for (int m = 0; m < M; ++m) xr[m] = xs[off + m];
#pragma unroll
      for (int dx = 0; dx < K; ++dx) {
        if (rho < K) {
#pragma unroll
          for (int j = 0; j < 4; ++j)
            dw[rho * K + dx] = __builtin_fmaf(g0[j], xr[j + dx], dw[rho * K + dx]);
        }
        if (rho >= 1) {
#pragma unroll
          for (int j = 0; j < 4; ++j)
            dw[(rho - 1) * K + dx] = __builtin_fmaf(g1[j], xr[j + dx], dw[(rho - 1) * K + dx]);
        }
      }
    }
    cur = nxt;
  }

  const int lane = tid & 63, wave = tid >> 6;
#pragma unroll
  for (int k = 0; k < KK; ++k) {
    const float v = wave_sum(dw[k]);
    if (lane == 0) red[wave * KK + k] = v;
  }
  __syncthreads();
  for (int k = tid; k < KK; k += 256)
    a.partials[(long long)blockIdx.x * KK + k] =
        ((red[k] + red[KK + k]) + red[2 * KK + k]) + red[3 * KK + k];
}

// dw by row streaming: a wave owns a strip of 62 lanes x 4 columns of dY and
// a segment of S rows; lanes 0 and 63 only supply the X halo.  The wave walks
// the X rows the segment's windows touch (S + 2R rows) and keeps the K dY
// rows each X row pairs with in a shifting register ring: X row iy meets dY
// row iy + R - dy through tap row dy.  The X window of a lane comes from its
// neighbours through DPP shifts.  Blocks are single waves.  Interior waves
// stage their X and dY rows PF rows ahead in a wave-private LDS ring with
// direct global->LDS loads (waits counted per wave, no barrier); waves that
// touch the image border read X through the boundary map (per-lane column
// tables, scalar row remap) one row ahead in registers.
template <int R, int PF>
__global__ __launch_bounds__(64) void conv_dw_stream_kernel(const ConvGradArgs a, int strips,
                                                            int segs, int total_waves, int S) {
  constexpr int K = 2 * R + 1, KK = K * K, V = 4;
  constexpr int HL = (R + V - 1) / V;
  constexpr int NW = V + 2 * R;
  constexpr int STRIP = (64 - 2 * HL) * V;
  constexpr int NSL = PF + 1;                          // LDS ring slots
  __shared__ __align__(16) float4 xs[NSL * 64];
  __shared__ __align__(16) float4 gls[NSL * 64];
  const int lane = threadIdx.x;
  float dw[KK];
#pragma unroll
  for (int t = 0; t < KK; ++t) dw[t] = 0.0f;

  for (int wv = blockIdx.x; wv < total_waves; wv += gridDim.x) {
    const int per_img = strips * segs;
    const int b = wv / per_img;
    const int rem = wv - b * per_img;
    const int strip = rem / segs, seg = rem - strip * segs;
    const float* X = a.x + b * a.x_plane;
    const float* G = a.gy + b * a.gy_plane;
    const int x_wave = strip * STRIP - HL * V;        // column of lane 0's pixel 0
    const int xl = x_wave + lane * V;
    const bool own = lane >= HL && lane < 64 - HL;
    const int y0 = seg * S;
    const int ny = min(S, a.h - y0);
    const int nx = ny + 2 * R;                         // X rows streamed
    const bool fast = a.vec_ok && x_wave >= 0 && x_wave + 64 * V <= a.w && y0 - R >= 0 &&
                      y0 + ny + R <= a.h;

    float ring[K][V];
#pragma unroll
    for (int d = 0; d < K; ++d)
#pragma unroll
      for (int e = 0; e < V; ++e) ring[d][e] = 0.0f;
    // X row k (absolute y0 - R + k) pairs with the ring of dY rows; the
    // newest ring entry is dY row y0 + k (zero past the segment)
    auto step = [&](const float (&cur)[V], const float (&gn)[V], int k) {
#pragma unroll
      for (int d = K - 1; d > 0; --d)
#pragma unroll
        for (int e = 0; e < V; ++e) ring[d][e] = ring[d - 1][e];
#pragma unroll
      for (int e = 0; e < V; ++e) ring[0][e] = gn[e];
      float win[NW];
#pragma unroll
      for (int wi = 0; wi < NW; ++wi) {
        const int rel = wi - R;
        const int dd = rel >= 0 ? rel / V : -((-rel + V - 1) / V);
        const int e = rel - dd * V;
        win[wi] = dd == 0 ? cur[e] : lane_shift(cur[e], dd);
      }
      // dY row y0 + k - dy exists for k - ny < dy <= k
      const int dmin = k - ny + 1, dmax = k;
#pragma unroll
      for (int dy = 0; dy < K; ++dy) {
        if (dy < dmin || dy > dmax) continue;
#pragma unroll
        for (int dx = 0; dx < K; ++dx)
#pragma unroll
          for (int j = 0; j < V; ++j)
            dw[dy * K + dx] = __builtin_fmaf(ring[dy][j], win[j + dx], dw[dy * K + dx]);
      }
    };

    if (fast) {
      // chunk k: X row y0 - R + k and dY row y0 + min(k, ny - 1) into slot
      // k % NSL; the rows past the stream are clamped (their values unused)
      auto issue = [&](int k) {
        const int slot = k % NSL;
        const int kx = min(k, nx - 1), kg = min(k, ny - 1);
        __builtin_amdgcn_global_load_lds(
            static_cast<const void*>(X + (long long)(y0 - R + kx) * a.x_stride + xl),
            (__attribute__((address_space(3))) void*)(xs + slot * 64), 16, 0, 0);
        __builtin_amdgcn_global_load_lds(
            static_cast<const void*>(G + (long long)(y0 + kg) * a.gy_stride + xl),
            (__attribute__((address_space(3))) void*)(gls + slot * 64), 16, 0, 0);
      };
#pragma unroll
      for (int k = 0; k < PF; ++k) issue(k);
      for (int k = 0; k < nx; ++k) {
        // rows k+1 .. k+PF-1 (2 loads each) may still be in flight
        asm volatile("s_waitcnt vmcnt(%0)" ::"n"(2 * (PF - 1)) : "memory");
        const float4 xq = xs[(k % NSL) * 64 + lane];
        const float4 gq = gls[(k % NSL) * 64 + lane];
        const bool gv = own && k < ny;
        const float cur[V] = {xq.x, xq.y, xq.z, xq.w};
        const float gn[V] = {gv ? gq.x : 0.0f, gv ? gq.y : 0.0f, gv ? gq.z : 0.0f,
                             gv ? gq.w : 0.0f};
        issue(k + PF);   // into the slot row k-1 used
        step(cur, gn, k);
      }
      asm volatile("s_waitcnt vmcnt(0)" ::: "memory");   // the ring is reused next
    } else {
      Geom g;
      g.in = X;
      g.in_stride = a.x_stride;
      g.lo_x = 0;
      g.hi_x = a.w;
      g.lo_y = 0;
      g.hi_y = a.h;
      g.in_imgw = a.w;
      g.in_imgh = a.h;
      g.boundary = a.boundary;
      g.const_f = a.cval;
      g.const_i = 0;
      g.in_dtype = HIPACC_F32;
      const LaneCols<V> cols = lane_cols<V>(g, xl);
      auto load_x = [&](float (&v)[V], int k) {
        bool rc;
        const int y = row_src(g, y0 - R + k, rc);
        const float* row = X + (long long)y * a.x_stride;
#pragma unroll
        for (int e = 0; e < V; ++e) {
          const float t = row[cols.x[e]];
          v[e] = (rc || cols.cst[e]) ? a.cval : t;
        }
      };
      auto load_g = [&](float (&v)[V], int k) {
#pragma unroll
        for (int e = 0; e < V; ++e) v[e] = 0.0f;
        if (k >= ny || !own) return;
        const float* row = G + (long long)(y0 + k) * a.gy_stride;
#pragma unroll
        for (int e = 0; e < V; ++e)
          if (xl + e < a.w) v[e] = row[xl + e];
      };
      float xn[V], gn[V];
      load_x(xn, 0);
      load_g(gn, 0);
      for (int k = 0; k < nx; ++k) {
        float cur[V], gk[V];
#pragma unroll
        for (int e = 0; e < V; ++e) {
          cur[e] = xn[e];
          gk[e] = gn[e];
        }
        if (k + 1 < nx) {
          load_x(xn, k + 1);
          load_g(gn, k + 1);
        }
        step(cur, gk, k);
      }
    }
  }

#pragma unroll
  for (int t = 0; t < KK; ++t) {
    const float v = wave_sum(dw[t]);
    if (lane == 0) a.partials[(long long)blockIdx.x * KK + t] = v;
  }
}

// Any mask shape: block (chunk, tap) sums dY(p) * X_bh(p + t - r) over the
// pixels p = chunk*256 + tid + k*nchunks*256 (all images).
__global__ __launch_bounds__(256) void conv_dw_generic(const ConvGradArgs a) {
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

int launch_conv_dw(const ConvGradArgs& a0, float* dw, hipStream_t s, PartialSum* sum) {
  ConvGradArgs a = a0;
  const int kk = a.kx * a.ky;
  int nblocks;
  if (a.kx == a.ky && (a.kx == 3 || a.kx == 5 || a.kx == 7 || a.kx == 9) && a.dw_variant != kVarTile) {
    constexpr int PF = 8;
    const int strip = (64 - 2) * 4;   // HL = 1 for R <= 4
    const int strips = ceil_div(a.w, strip);
    // Segment heights such that all segments run in one round of the
    // kGradBlocks resident waves (2 per SIMD at this kernel's registers): a
    // second, partial round would double the time of the waves that take
    // it (16 x 1024^2, 9x9: 32-row segments, 2560 waves, 111 us).
    const long long per_seg_row = (long long)strips * a.batch;
    int segs_fit = (int)std::max(1LL, (long long)kGradBlocks / per_seg_row);
    const int S = std::max(ceil_div(a.h, segs_fit), 8);
    const int segs = ceil_div(a.h, S);
    const int waves = strips * segs * a.batch;
    nblocks = min(waves, kGradBlocks);
    switch (a.kx) {
      case 3: hipLaunchKernelGGL((conv_dw_stream_kernel<1, PF>), dim3(nblocks), dim3(64), 0, s, a, strips, segs, waves, S); break;
      case 5: hipLaunchKernelGGL((conv_dw_stream_kernel<2, PF>), dim3(nblocks), dim3(64), 0, s, a, strips, segs, waves, S); break;
      case 7: hipLaunchKernelGGL((conv_dw_stream_kernel<3, PF>), dim3(nblocks), dim3(64), 0, s, a, strips, segs, waves, S); break;
      default: hipLaunchKernelGGL((conv_dw_stream_kernel<4, PF>), dim3(nblocks), dim3(64), 0, s, a, strips, segs, waves, S); break;
    }
  } else if (a.kx == a.ky && (a.kx == 3 || a.kx == 5 || a.kx == 7 || a.kx == 9)) {
    a.tiles_x = ceil_div(a.w, 64);
    a.tiles_y = ceil_div(a.h, 32);
    a.ntiles = a.tiles_x * a.tiles_y * a.batch;
    nblocks = a.ntiles < kGradBlocks ? a.ntiles : kGradBlocks;
    const int per = ceil_div(a.ntiles, nblocks);
    nblocks = ceil_div(a.ntiles, per);
    switch (a.kx) {
      case 3: hipLaunchKernelGGL(conv_dw_kernel<1>, dim3(nblocks), dim3(256), 0, s, a, per); break;
      case 5: hipLaunchKernelGGL(conv_dw_kernel<2>, dim3(nblocks), dim3(256), 0, s, a, per); break;
      case 7: hipLaunchKernelGGL(conv_dw_kernel<3>, dim3(nblocks), dim3(256), 0, s, a, per); break;
      default: hipLaunchKernelGGL(conv_dw_kernel<4>, dim3(nblocks), dim3(256), 0, s, a, per); break;
    }
  } else {
    const long long total = (long long)a.w * a.h * a.batch;
    const long long nc = (total + 256 * 16 - 1) / (256 * 16);
    nblocks = (int)(nc < 256 ? nc : 256);
    hipLaunchKernelGGL(conv_dw_generic, dim3(nblocks, kk), dim3(256), 0, s, a);
  }
  sum->partials = a.partials;
  sum->nblocks = nblocks;
  sum->kk = kk;
  sum->scale = 1.0f;
  sum->out = dw;
  return hip_status(hipGetLastError(), "conv_dw launch");
}

}  // namespace hh
