// bilateral.hip -- the Bilateral_Filter sample's operator
// (samples-public/3_Preprocessing/Bilateral_Filter/src/main.cpp:63-77).
//
// Same tile staging and row streaming as the local SUM kernels; per tap the
// range weight and the spatial weight are combined in the log2 domain,
//     s = exp2(-c_r*log2(e) * diff^2 + log2(spatial)),
// so a tap costs one subtract, one multiply, one FMA, one v_exp_f32 and two
// accumulations.  This evaluates expf(-c_r*diff^2)*spatial with a different
// rounding; the tolerance it meets is stated in the tests.
#include "hh_tile.hpp"

namespace hh {

template <typename TI, typename TO, int KS, int PPT>
__global__ __launch_bounds__(256) void bilateral_kernel(const BilatArgs a) {
  constexpr int VEC = 4;  // 4 pixels per lane for both float and uchar images
  constexpr int R = KS / 2;
  constexpr int HV = (R + VEC - 1) / VEC;
  constexpr int NV = 64 + 2 * HV;
  constexpr int LW = NV * VEC;
  constexpr int TW = 64 * VEC, TH = 4 * PPT;
  constexpr int LH = TH + 2 * R;
  constexpr int NW = 1 + 2 * HV;
  __shared__ __align__(16) float tile[LH * LW];

  const Geom& g = a.g;
  const int tx0 = blockIdx.x * TW;
  const int ty0 = blockIdx.y * TH;
  const TI* in = static_cast<const TI*>(g.in) + (long long)blockIdx.z * g.in_plane;
  TO* out = static_cast<TO*>(g.out) + (long long)blockIdx.z * g.out_plane;
  const int ix0 = g.in_ox + tx0 - HV * VEC, iy0 = g.in_oy + ty0 - R;
  if constexpr (std::is_same<TI, float>::value) {
    stage_tile<float, VEC, NV, LH>(tile, in, g, ix0, iy0, a.in_vec_ok);
  } else {
    // uchar images are widened to float while staging (per element).
    const bool cmode = g.boundary == HIPACC_BOUNDARY_CONSTANT;
    int lx, hx, ly, hy;
    raw_region(g, lx, hx, ly, hy);
    const bool interior = ix0 >= lx && ix0 + LW <= hx && iy0 >= ly && iy0 + LH <= hy;
    for (int i = threadIdx.x; i < LH * LW; i += 256) {
      const int r = i / LW, c = i - r * LW;
      int x = ix0 + c, y = iy0 + r;
      float v;
      if (!interior && cmode && (x < g.lo_x || x >= g.hi_x || y < g.lo_y || y >= g.hi_y)) {
        v = (float)(TI)g.const_i;
      } else {
        if (!interior && !cmode) {
          x = bh_map(g.boundary, x, g.lo_x, g.hi_x, g.in_imgw);
          y = bh_map(g.boundary, y, g.lo_y, g.hi_y, g.in_imgh);
        }
        v = (float)in[(long long)y * g.in_stride + x];
      }
      tile[i] = v;
    }
  }
  __syncthreads();

  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const float* base = tile + (wave * PPT) * LW + lane * VEC;
  float ctr[PPT][VEC], num[PPT][VEC], den[PPT][VEC];
#pragma unroll
  for (int o = 0; o < PPT; ++o) {
    const float* c = base + (o + R) * LW + HV * VEC;
#pragma unroll
    for (int v = 0; v < VEC; ++v) {
      ctr[o][v] = c[v];
      num[o][v] = 0.0f;
      den[o][v] = 0.0f;
    }
  }
  const float k2 = a.k2;
#pragma unroll
  for (int ir = 0; ir < PPT + 2 * R; ++ir) {
    float win[NW * VEC];
    load_window<float, float, VEC, NW>(win, base + ir * LW);
#pragma unroll
    for (int o = 0; o < PPT; ++o) {
      const int dy = ir - o;
      if (dy < 0 || dy >= KS) continue;
#pragma unroll
      for (int dx = 0; dx < KS; ++dx) {
        const float l = a.lsp[dy * KS + dx];
#pragma unroll
        for (int v = 0; v < VEC; ++v) {
          const float x = win[HV * VEC + v + dx - R];
          const float d = x - ctr[o][v];
          const float s = __builtin_amdgcn_exp2f(__builtin_fmaf(k2, d * d, l));
          den[o][v] += s;
          num[o][v] = __builtin_fmaf(s, x, num[o][v]);
        }
      }
    }
  }
  const int ox = tx0 + lane * VEC;
#pragma unroll
  for (int o = 0; o < PPT; ++o) {
    float r[VEC];
#pragma unroll
    for (int v = 0; v < VEC; ++v) r[v] = num[o][v] / den[o][v];
    store_row<TO, float, VEC>(out, g, a.out_vec_ok, ox, ty0 + wave * PPT + o, r);
  }
  if constexpr (std::is_same<TO, float>::value) {
    if (a.norm) {
      float* nrm = a.norm + (long long)blockIdx.z * a.norm_plane;
#pragma unroll
      for (int o = 0; o < PPT; ++o) {
        const int y = ty0 + wave * PPT + o;
        if (y >= g.is_h) continue;
#pragma unroll
        for (int v = 0; v < VEC; ++v)
          if (ox + v < g.is_w) nrm[(long long)y * a.norm_stride + ox + v] = den[o][v];
      }
    }
  }
}

template <typename TI, typename TO, int KS>
static int launch_bilat_t(const BilatArgs& a, hipStream_t s) {
  constexpr int PPT = KS >= 9 ? 2 : 4;
  dim3 grid(ceil_div(a.g.is_w, 256), ceil_div(a.g.is_h, 4 * PPT), a.g.batch);
  hipLaunchKernelGGL((bilateral_kernel<TI, TO, KS, PPT>), grid, dim3(256), 0, s, a);
  return hip_status(hipGetLastError(), "bilateral_kernel launch");
}

int launch_bilateral(const BilatArgs& a, hipStream_t s) {
  const bool f = a.g.in_dtype == HIPACC_F32;
#define HH_BIL(K)                                                              \
  case K:                                                                      \
    return f ? launch_bilat_t<float, float, K>(a, s) : launch_bilat_t<uint8_t, uint8_t, K>(a, s);
  switch (a.ksize) {
    HH_BIL(3)
    HH_BIL(5)
    HH_BIL(7)
    HH_BIL(9)
    HH_BIL(11)
    HH_BIL(13)
    HH_BIL(15)
  }
#undef HH_BIL
  return set_error(HIPACC_ERR_UNSUPPORTED, "bilateral: ksize %d", a.ksize);
}

}  // namespace hh
