// grad_bilat.hip -- backward of the bilateral filter with a learnable range
// sigma.  Forward (per output pixel p, taps t at offsets o_t, v = X_bh(p+o_t)):
//     s = exp(-c (v - X(p))^2) * sp[t],  D(p) = SUM s,  Y(p) = SUM s*v / D(p)
// With a(p) = dY(p)/D(p), d = v - X(q):
//   centre term of q      2c a(q) SUM_t s_{q,t} d (v - Y(q))
//   neighbour terms of q  SUM a(p) s_{p,t} [1 + 2c d' (X(q) - Y(p))] over the
//                         (p, t) whose tap reads q, d' = X(p) - X(q)
//   dL/dc                 -SUM_q a(q) SUM_t s_{q,t} d^2 (v - Y(q))
// Away from the borders the neighbour p = q + o reads q through the opposite
// tap -o, and exp(-c d^2) is the same for the pair, so one exponential per
// (q, o) serves both terms.  Pixels within r of an edge also receive the
// neighbour terms of the halo positions the boundary mode folds onto them:
// those are evaluated once per halo (ring) position and added afterwards.
#include "hh_grad.hpp"

namespace hh {

// Neighbour terms landing on ring position i (outside the image, within R
// of it), read by the taps of the in-image pixels p = i - o_t; the value
// read there is X(q), q = bh(i).  K threads per position, one per tap row:
// each sums its row's taps inside the image, and the position's first
// thread adds the K row sums in row order.
template <int R>
__global__ __launch_bounds__(256) void bilat_ring_kernel(const BilatGradArgs a) {
  constexpr int K = 2 * R + 1, P = 256 / K;   // positions per block
  __shared__ float rows[P * K];
  const int t = threadIdx.x, pl = t / K, ty = t - pl * K;
  const int j = blockIdx.x * P + pl;
  const int b = blockIdx.y;
  const bool live = pl < P && j < (int)a.rsize;
  float acc = 0.0f;
  int ix = 0, iy = 0;
  if (live) {
    ring_pos<int>(a.w, a.h, R, R, j, ix, iy);
    const int y = iy - (ty - R);
    if (y >= 0 && y < a.h) {
      const int qx = bh_map(a.boundary, ix, 0, a.w, a.w);
      const int qy = bh_map(a.boundary, iy, 0, a.h, a.h);
      const float* X = a.x + b * a.x_plane;
      const float* Y = a.y + b * a.y_plane;
      const float* D = a.d + b * a.d_plane;
      const float* G = a.gy + b * a.gy_plane;
      const float xq = X[(long long)qy * a.x_stride + qx];
      for (int tx = 0; tx < K; ++tx) {
        const int x = ix - (tx - R);
        if (x < 0 || x >= a.w) continue;
        const float xp = X[(long long)y * a.x_stride + x];
        const float yp = Y[(long long)y * a.y_stride + x];
        const float ap = G[(long long)y * a.gy_stride + x] / D[(long long)y * a.d_stride + x];
        const float dd = xp - xq;
        const float sv = __builtin_amdgcn_exp2f(a.k2 * (dd * dd)) * a.sp[ty * K + tx];
        const float w = ap * sv;
        acc += w + 2.0f * a.c * (w * dd) * (xq - yp);
      }
    }
  }
  if (pl < P) rows[t] = acc;
  __syncthreads();
  if (live && ty == 0) {
    float sum = rows[pl * K];
#pragma unroll
    for (int r = 1; r < K; ++r) sum += rows[pl * K + r];
    a.ring[b * a.rsize + j] = sum;
  }
}

// 64 x 16 pixel tiles, 4 adjacent pixels per lane.  LDS holds, with a halo of
// r rows and RH columns: X_bh (CONSTANT: the constant), a = dY/D (0 outside
// the image) and Y.
// SYM: the spatial table is point-symmetric (sp[t] == sp[K*K-1-t], as the
// sigma_d table is), so the weight of a pixel pair is one exponential
// exp2(k2*d^2 + log2 sp) for both its centre and its neighbour term.
template <int R, bool DX, bool SYM>
__global__ __launch_bounds__(256) void bilat_grad_kernel(const BilatGradArgs a) {
  constexpr int K = 2 * R + 1, KK = K * K;
  constexpr int TW = 64, TH = 16;
  constexpr int RH = (R + 3) & ~3;
  constexpr int LW = TW + 2 * RH, LH = TH + 2 * R;
  constexpr int M = 4 + 2 * R;
  __shared__ __align__(16) float xs[LH * LW];
  __shared__ __align__(16) float as[LH * LW];
  __shared__ __align__(16) float ys[LH * LW];
  __shared__ float red[4];

  const int tid = threadIdx.x, tr = tid >> 4, tc = (tid & 15) * 4;
  const int per_plane = a.tiles_x * a.tiles_y;
  const float k2 = a.k2, c2 = 2.0f * a.c;
  float dc = 0.0f;

  for (int tile = blockIdx.x; tile < a.ntiles; tile += gridDim.x) {
    const int b = tile / per_plane, rem = tile - b * per_plane;
    const int y0 = (rem / a.tiles_x) * TH, x0 = (rem % a.tiles_x) * TW;
    const float* X = a.x + b * a.x_plane;
    const float* Y = a.y + b * a.y_plane;
    const float* D = a.d + b * a.d_plane;
    const float* G = a.gy + b * a.gy_plane;
    const bool interior =
        a.vec_ok && x0 >= RH && x0 + TW + RH <= a.w && y0 >= R && y0 + TH + R <= a.h;
    __syncthreads();
    if (interior) {
      constexpr int LW4 = LW / 4, N4 = LH * LW4;
      for (int i = tid; i < N4; i += 256) {
        const int r = i / LW4, c = (i - r * LW4) * 4;
        const long long gx = x0 - RH + c;
        const int gy = y0 - R + r;
        const float4 xv = *reinterpret_cast<const float4*>(X + (long long)gy * a.x_stride + gx);
        const float4 yv = *reinterpret_cast<const float4*>(Y + (long long)gy * a.y_stride + gx);
        const float4 dv = *reinterpret_cast<const float4*>(D + (long long)gy * a.d_stride + gx);
        const float4 gv = *reinterpret_cast<const float4*>(G + (long long)gy * a.gy_stride + gx);
        *reinterpret_cast<float4*>(xs + r * LW + c) = xv;
        *reinterpret_cast<float4*>(ys + r * LW + c) = yv;
        *reinterpret_cast<float4*>(as + r * LW + c) =
            make_float4(gv.x / dv.x, gv.y / dv.y, gv.z / dv.z, gv.w / dv.w);
      }
    } else {
      for (int i = tid; i < LH * LW; i += 256) {
        const int r = i / LW, c = i - r * LW;
        const int y = y0 - R + r, x = x0 - RH + c;
        const bool inside = x >= 0 && x < a.w && y >= 0 && y < a.h;
        float xv, av = 0.0f, yv = 0.0f;
        if (inside) {
          xv = X[(long long)y * a.x_stride + x];
          yv = Y[(long long)y * a.y_stride + x];
          av = G[(long long)y * a.gy_stride + x] / D[(long long)y * a.d_stride + x];
        } else if (a.boundary == HIPACC_BOUNDARY_CONSTANT) {
          xv = a.cval;
        } else {
          xv = X[(long long)bh_map(a.boundary, y, 0, a.h, a.h) * a.x_stride +
                 bh_map(a.boundary, x, 0, a.w, a.w)];
        }
        xs[i] = xv;
        as[i] = av;
        ys[i] = yv;
      }
    }
    __syncthreads();

    float xq[4], cq[4], aq[4];
    float accc[4] = {0.0f, 0.0f, 0.0f, 0.0f}, accs[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    float accn[4] = {0.0f, 0.0f, 0.0f, 0.0f}, accm[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int o = (tr + R) * LW + tc + RH + j;
      xq[j] = xs[o];
      cq[j] = xs[o] - ys[o];   // X(q) - Y(q)
      aq[j] = as[o];
    }
    // One tap row per iteration (not unrolled): the three row windows stay
    // within the register budget of 4 waves per SIMD.
#pragma unroll 1
    for (int dy = 0; dy < K; ++dy) {
      // row windows read as the 1 + RH/2 aligned 16-byte words that hold
      // them (conflict-free: a wave's 16-lane row groups read consecutive
      // words; single floats at a 4-word lane stride hit 8 banks only)
      constexpr int NK = 1 + RH / 2;
      float xf4[4 * NK], af4[DX ? 4 * NK : 1], yf4[DX ? 4 * NK : 1];
      const int base = (tr + dy) * LW + tc;
#pragma unroll
      for (int k = 0; k < NK; ++k) {
        const float4 q = *reinterpret_cast<const float4*>(xs + base + 4 * k);
        xf4[4 * k] = q.x;
        xf4[4 * k + 1] = q.y;
        xf4[4 * k + 2] = q.z;
        xf4[4 * k + 3] = q.w;
        if constexpr (DX) {
          const float4 qa = *reinterpret_cast<const float4*>(as + base + 4 * k);
          const float4 qy = *reinterpret_cast<const float4*>(ys + base + 4 * k);
          af4[4 * k] = qa.x;
          af4[4 * k + 1] = qa.y;
          af4[4 * k + 2] = qa.z;
          af4[4 * k + 3] = qa.w;
          yf4[4 * k] = qy.x;
          yf4[4 * k + 1] = qy.y;
          yf4[4 * k + 2] = qy.z;
          yf4[4 * k + 3] = qy.w;
        }
      }
      float xr[M], ar[DX ? M : 1], zr[DX ? M : 1];
#pragma unroll
      for (int m = 0; m < M; ++m) {
        xr[m] = xf4[RH - R + m];
        if constexpr (DX) {
          ar[m] = af4[RH - R + m];
          zr[m] = xr[m] - yf4[RH - R + m];   // X(p) - Y(p)
        }
      }
#pragma unroll
      for (int dx = 0; dx < K; ++dx) {
        const float spc = a.sp[dy * K + dx];           // tap o of q
        const float spn = a.sp[KK - 1 - (dy * K + dx)]; // tap -o of p = q + o
        const float lsp = a.lsp[dy * K + dx];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const float v = xr[j + dx];
          const float d = v - xq[j];
          const float dd = d * d;
          float sc, sn;   // s of the pair for the centre and the neighbour term
          if constexpr (SYM) {
            sc = sn = __builtin_amdgcn_exp2f(__builtin_fmaf(k2, dd, lsp));
          } else {
            const float e = __builtin_amdgcn_exp2f(k2 * dd);
            sc = e * spc;
            sn = e * spn;
          }
          // d (X(p) - Y(q)) = d (d + X(q) - Y(q)), and d (X(q) - Y(p)) =
          // d ((X(p) - Y(p)) - d): one FMA each from d^2
          const float u = sc * __builtin_fmaf(d, cq[j], dd);
          accc[j] += u;
          accs[j] = __builtin_fmaf(u, d, accs[j]);
          if constexpr (DX) {
            const float w = ar[j + dx] * sn;
            accn[j] += w;
            accm[j] = __builtin_fmaf(w, __builtin_fmaf(d, zr[j + dx], -dd), accm[j]);
          }
        }
      }
    }

    const int qy = y0 + tr;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int qx = x0 + tc + j;
      if (qy < a.h && qx < a.w) dc -= aq[j] * accs[j];
    }
    if constexpr (DX) {
      if (qy < a.h) {
        float* GX = a.gx + b * a.gx_plane + (long long)qy * a.gx_stride;
        float v[4];
        // neighbour term: SUM w [1 + 2c (X(p) - X(q)) (X(q) - Y(p))], d = X(p) - X(q)
#pragma unroll
        for (int j = 0; j < 4; ++j) v[j] = c2 * aq[j] * accc[j] + (accn[j] + c2 * accm[j]);
        if (interior) {
          *reinterpret_cast<float4*>(GX + x0 + tc) = make_float4(v[0], v[1], v[2], v[3]);
        } else {
#pragma unroll
          for (int j = 0; j < 4; ++j)
            if (x0 + tc + j < a.w) GX[x0 + tc + j] = v[j];
        }
      }
    }
  }

  dc = wave_sum(dc);
  if ((tid & 63) == 0) red[tid >> 6] = dc;
  __syncthreads();
  if (tid == 0) a.partials[blockIdx.x] = ((red[0] + red[1]) + red[2]) + red[3];
}

template <int R>
static int launch_bilat_ring(const BilatGradArgs& a, hipStream_t s) {
  constexpr int P = 256 / (2 * R + 1);
  hipLaunchKernelGGL(bilat_ring_kernel<R>, dim3((unsigned)((a.rsize + P - 1) / P), a.batch),
                     dim3(256), 0, s, a);
  return hip_status(hipGetLastError(), "bilat_ring_kernel launch");
}

template <int R>
static int launch_bg(const BilatGradArgs& a, int nblocks, hipStream_t s) {
  if (a.sym) {
    if (a.gx)
      hipLaunchKernelGGL((bilat_grad_kernel<R, true, true>), dim3(nblocks), dim3(256), 0, s, a);
    else
      hipLaunchKernelGGL((bilat_grad_kernel<R, false, true>), dim3(nblocks), dim3(256), 0, s, a);
  } else if (a.gx) {
    hipLaunchKernelGGL((bilat_grad_kernel<R, true, false>), dim3(nblocks), dim3(256), 0, s, a);
  } else {
    hipLaunchKernelGGL((bilat_grad_kernel<R, false, false>), dim3(nblocks), dim3(256), 0, s, a);
  }
  return hip_status(hipGetLastError(), "bilat_grad_kernel launch");
}

int launch_bilateral_grad(const BilatGradArgs& a0, float* dsigma, float sigma_r, float* ring,
                          hipStream_t s) {
  BilatGradArgs a = a0;
  if (!a.gx && !dsigma) return HIPACC_OK;
  a.ring = a.gx && a.boundary != HIPACC_BOUNDARY_CONSTANT ? ring : nullptr;
  a.rsize = ring_size(a.w, a.h, a.ksize / 2, a.ksize / 2);
  if (a.rsize == 0) a.ring = nullptr;
  if (a.ring && a.rsize + 256 >= (1LL << 31))
    return set_error(HIPACC_ERR_UNSUPPORTED, "bilateral_grad: image too large for the halo ring");
  a.tiles_x = ceil_div(a.w, 64);
  a.tiles_y = ceil_div(a.h, 16);
  a.ntiles = a.tiles_x * a.tiles_y * a.batch;
  const int nblocks = a.ntiles < kGradBlocks ? a.ntiles : kGradBlocks;
  if (a.ksize > 15) return set_error(HIPACC_ERR_UNSUPPORTED, "bilateral_grad: ksize %d", a.ksize);
  int st;
  switch (a.ksize) {
    case 1: st = launch_bg<0>(a, nblocks, s); break;
    case 3: st = launch_bg<1>(a, nblocks, s); break;
    case 5: st = launch_bg<2>(a, nblocks, s); break;
    case 7: st = launch_bg<3>(a, nblocks, s); break;
    case 9: st = launch_bg<4>(a, nblocks, s); break;
    case 11: st = launch_bg<5>(a, nblocks, s); break;
    case 13: st = launch_bg<6>(a, nblocks, s); break;
    case 15: st = launch_bg<7>(a, nblocks, s); break;
    default: return set_error(HIPACC_ERR_UNSUPPORTED, "bilateral_grad: ksize %d", a.ksize);
  }
  if (st == HIPACC_OK && a.ring) {
    switch (a.ksize) {
      case 3: st = launch_bilat_ring<1>(a, s); break;
      case 5: st = launch_bilat_ring<2>(a, s); break;
      case 7: st = launch_bilat_ring<3>(a, s); break;
      case 9: st = launch_bilat_ring<4>(a, s); break;
      case 11: st = launch_bilat_ring<5>(a, s); break;
      case 13: st = launch_bilat_ring<6>(a, s); break;
      default: st = launch_bilat_ring<7>(a, s); break;
    }
  }
  // dsigma = -1/sigma^3 * SUM_b partials[b]: with the fold when there is one
  PartialSum ps{a.partials, nblocks, 1, (float)(-1.0 / ((double)sigma_r * sigma_r * sigma_r)),
                dsigma};
  if (st == HIPACC_OK && a.ring)
    st = launch_fold_add(a.gx, a.gx_plane, a.gx_stride, a.w, a.h, a.batch, a.ksize / 2,
                         a.ksize / 2, a.boundary, a.ring, s, dsigma ? &ps : nullptr);
  else if (st == HIPACC_OK && dsigma)
    st = launch_partial_sum(ps, s);
  return st;
}

}  // namespace hh
