// bilateral.hip -- the Bilateral_Filter sample's operator
// (samples-public/3_Preprocessing/Bilateral_Filter/src/main.cpp:63-77).
//
// Tile staging as the local SUM kernels; per tap the
// range weight and the spatial weight are combined in the log2 domain,
//     s = exp2(-c_r*log2(e) * diff^2 + log2(spatial)),
// so a tap costs one subtract, one multiply, one FMA, one v_exp_f32 and two
// accumulations.  This evaluates expf(-c_r*diff^2)*spatial with a different
// rounding; the tolerance it meets is stated in the tests.
#include "hh_tile.hpp"

namespace hh {

// 64 x 16 output tiles; each lane owns 4 adjacent pixels of one row.  The
// input tile with its halo is staged in LDS as float; the tap rows are
// walked one at a time (not unrolled), so a lane holds one row window of
// 4 + 2R values plus its 8 sums, whatever the mask size.
template <typename TI, typename TO, int KS>
__global__ __launch_bounds__(256) void bilateral_kernel(const BilatArgs a) {
  constexpr int VEC = 4;
  constexpr int R = KS / 2;
  constexpr int HV = (R + VEC - 1) / VEC;   // halo in 4-pixel groups
  constexpr int TW = 64, TH = 16;
  constexpr int NV = TW / VEC + 2 * HV;
  constexpr int LW = NV * VEC, LH = TH + 2 * R;
  constexpr int M = VEC + 2 * R;
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

  const int tr = threadIdx.x >> 4, tc = (threadIdx.x & 15) * VEC;
  float ctr[VEC], num[VEC], den[VEC];
#pragma unroll
  for (int v = 0; v < VEC; ++v) {
    ctr[v] = tile[(tr + R) * LW + HV * VEC + tc + v];
    num[v] = 0.0f;
    den[v] = 0.0f;
  }
  const float k2 = a.k2;
#pragma unroll 1
  for (int dy = 0; dy < KS; ++dy) {
    // the row window, read as the 1 + 2 HV aligned 16-byte words that hold
    // it: a wave's 16-lane row groups read consecutive words (no LDS bank
    // conflicts; single floats at a 4-word lane stride hit 8 banks only)
    constexpr int NK = 1 + 2 * HV;
    float flat[4 * NK];
    const float4* row4 = reinterpret_cast<const float4*>(tile + (tr + dy) * LW + tc);
#pragma unroll
    for (int k = 0; k < NK; ++k) {
      const float4 q = row4[k];
      flat[4 * k] = q.x;
      flat[4 * k + 1] = q.y;
      flat[4 * k + 2] = q.z;
      flat[4 * k + 3] = q.w;
    }
    float win[M];
#pragma unroll
    for (int m = 0; m < M; ++m) win[m] = flat[HV * VEC - R + m];
#pragma unroll
    for (int dx = 0; dx < KS; ++dx) {
      const float l = a.lsp[dy * KS + dx];
#pragma unroll
      for (int v = 0; v < VEC; ++v) {
        const float x = win[v + dx];
        const float d = x - ctr[v];
        const float sw = __builtin_amdgcn_exp2f(__builtin_fmaf(k2, d * d, l));
        den[v] += sw;
        num[v] = __builtin_fmaf(sw, x, num[v]);
      }
    }
  }
  float r[VEC];
#pragma unroll
  for (int v = 0; v < VEC; ++v) r[v] = num[v] / den[v];
  const int ox = tx0 + tc, oy = ty0 + tr;
  store_row<TO, float, VEC>(out, g, a.out_vec_ok, ox, oy, r);
  if constexpr (std::is_same<TO, float>::value) {
    if (a.norm && oy < g.is_h) {
      float* nrm = a.norm + (long long)blockIdx.z * a.norm_plane + (long long)oy * a.norm_stride;
#pragma unroll
      for (int v = 0; v < VEC; ++v)
        if (ox + v < g.is_w) nrm[ox + v] = den[v];
    }
  }
}

template <typename TI, typename TO, int KS>
static int launch_bilat_t(const BilatArgs& a, hipStream_t s) {
  dim3 grid(ceil_div(a.g.is_w, 64), ceil_div(a.g.is_h, 16), a.g.batch);
  hipLaunchKernelGGL((bilateral_kernel<TI, TO, KS>), grid, dim3(256), 0, s, a);
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
