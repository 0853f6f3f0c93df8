// local_vec4.hip -- local operators on 4-wide vector pixels (uchar4 / float4
// / short4 / int4 in, any vector type out): Kernel::convolve and
// Kernel::reduce(Domain) of the RGBA samples (e.g. samples-public/
// 1_Local_Operators/Gaussian_Blur_RGBA, Box_Blur_RGBA, Erode_RGBA,
// 3_Preprocessing/Sobel_RGBA).  The DSL's vector arithmetic is componentwise,
// so each channel is the scalar operator of local.hip: first active tap
// seeds, taps in row-major order, the same accumulator arithmetic and output
// conversion.
//
// 64 x 16 output tiles; the input tile and its halo are staged in LDS as
// four accumulator values per pixel (16 bytes: a wave's lanes read
// consecutive 16-byte words, free of bank conflicts), with the boundary
// mode applied while staging.  Each lane computes a column of 4 output
// pixels.
#include "hh_local.hpp"

namespace hh {

template <typename ACC> struct alignas(16) Px4 { ACC c[4]; };

template <typename ACC, int KW>
__global__ __launch_bounds__(256) void local_vec4_kernel(const LocalArgs a) {
  constexpr bool IS_F = std::is_same<ACC, float>::value;
  constexpr int TW = 64, TH = 16;
  extern __shared__ __align__(16) unsigned char vec4_lds[];
  Px4<ACC>* tile = reinterpret_cast<Px4<ACC>*>(vec4_lds);
  const Geom& g = a.g;
  const int mw = KW ? KW : a.mask_w, mh = KW ? KW : a.mask_h;
  const int rx = mw / 2, ry = mh / 2;
  const int LW = TW + mw - 1, LH = TH + mh - 1;
  const int x0 = blockIdx.x * TW, y0 = blockIdx.y * TH;
  const int ib = dtype_base(g.in_dtype), ob = dtype_base(g.out_dtype);
  const char* in = static_cast<const char*>(g.in) +
                   (long long)blockIdx.z * g.in_plane * dtype_bytes(g.in_dtype);
  char* out = static_cast<char*>(g.out) +
              (long long)blockIdx.z * g.out_plane * dtype_bytes(g.out_dtype);
  const bool cmode = g.boundary == HIPACC_BOUNDARY_CONSTANT;
  const ACC cv = IS_F ? (ACC)g.const_f : (ACC)g.const_i;
  for (int i = threadIdx.x; i < LW * LH; i += 256) {
    const int r = i / LW, c = i - r * LW;
    int xx = g.in_ox + x0 + c - rx, yy = g.in_oy + y0 + r - ry;
    Px4<ACC> v;
    bool cst = false;
    if (cmode) {
      cst = xx < g.lo_x || xx >= g.hi_x || yy < g.lo_y || yy >= g.hi_y;
    } else {
      xx = bh_map(g.boundary, xx, g.lo_x, g.hi_x, g.in_imgw);
      yy = bh_map(g.boundary, yy, g.lo_y, g.hi_y, g.in_imgh);
    }
    const long long idx = ((long long)yy * g.in_stride + xx) * 4;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if constexpr (IS_F) v.c[k] = cst ? cv : load_f(in, ib, idx + k);
      else v.c[k] = cst ? cv : load_i(in, ib, idx + k);
    }
    tile[i] = v;
  }
  __syncthreads();

  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int x = x0 + tx;
#pragma unroll 1
  for (int j = 0; j < 4; ++j) {
    const int yl = ty * 4 + j, y = y0 + yl;
    ACC acc[4] = {ACC(0), ACC(0), ACC(0), ACC(0)};
    bool first = true;
    for (int dy = 0; dy < mh; ++dy) {
#pragma unroll(KW ? KW : 1)
      for (int dx = 0; dx < mw; ++dx) {
        const int t = dy * mw + dx;
        if (!a.domain[t]) continue;
        const Px4<ACC> v = tile[(yl + dy) * LW + tx + dx];
        ACC w = ACC(1);
        if (a.use_weights) {
          if constexpr (IS_F) w = a.w.f[t];
          else w = a.w.i[t];
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          if (a.reduce == HIPACC_REDUCE_SUM && a.use_weights) {
            if (first) acc[k] = madd<false>(w, v.c[k], ACC(0));
            else if (a.fma) acc[k] = madd<true>(w, v.c[k], acc[k]);
            else acc[k] = madd<false>(w, v.c[k], acc[k]);
          } else {
            ACC tap = a.use_weights ? (ACC)(w * v.c[k]) : v.c[k];
            if constexpr (!IS_F)
              if (a.use_weights) tap = (ACC)((uint32_t)w * (uint32_t)v.c[k]);
            acc[k] = first ? tap : reduce_combine<ACC>(a.reduce, acc[k], tap);
          }
        }
        first = false;
      }
    }
    if (x < g.is_w && y < g.is_h) {
      const long long o = ((long long)(g.out_oy + y) * g.out_stride + g.out_ox + x) * 4;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        if constexpr (IS_F) store_f(out, ob, o + k, acc[k], g.cvt, g.post_div);
        else store_i(out, ob, o + k, acc[k], g.cvt, g.post_div);
      }
    }
  }
}

template <typename ACC, int KW>
static int launch_vec4_t(const LocalArgs& a, hipStream_t s) {
  const size_t smem = sizeof(Px4<ACC>) * (size_t)(64 + a.mask_w - 1) * (16 + a.mask_h - 1);
  dim3 grid(ceil_div(a.g.is_w, 64), ceil_div(a.g.is_h, 16), a.g.batch);
  hipLaunchKernelGGL((local_vec4_kernel<ACC, KW>), grid, dim3(256), smem, s, a);
  return hip_status(hipGetLastError(), "local_vec4_kernel launch");
}

int launch_local_vec4(const LocalArgs& a, int acc_is_float, hipStream_t s) {
  if (a.wdev) return set_error(HIPACC_ERR_UNSUPPORTED, "local: device weights on vector pixels");
  const bool sq = a.mask_w == a.mask_h;
#define HH_V4(K)                                                                          \
  if (sq && a.mask_w == K)                                                                \
    return acc_is_float ? launch_vec4_t<float, K>(a, s) : launch_vec4_t<int32_t, K>(a, s);
  HH_V4(3) HH_V4(5) HH_V4(7)
#undef HH_V4
  return acc_is_float ? launch_vec4_t<float, 0>(a, s) : launch_vec4_t<int32_t, 0>(a, s);
}

}  // namespace hh
