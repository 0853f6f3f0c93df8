// separable.hip -- row pass then column pass in one launch.
//
// The staged tile is the same as the 2D kernel's.  Each lane computes the row
// pass for its VEC columns on every staged row its outputs need and feeds it
// straight into the column-pass accumulators, so the intermediate image of
// the two-kernel DSL program never leaves registers.  For CONSTANT, rows
// outside the accessor window contribute the constant itself (the value the
// intermediate image's own boundary condition returns).
#include "hh_tile.hpp"

namespace hh {

template <typename TI, typename TO, typename ACC, int KX, int KY, int PPT, bool FMA, bool BOX>
__global__ __launch_bounds__(256) void separable_kernel(const SepArgs a) {
  constexpr int VEC = 16 / (int)sizeof(TI);
  constexpr int RX = KX / 2, RY = KY / 2;
  constexpr int HV = (RX + VEC - 1) / VEC;
  constexpr int NV = 64 + 2 * HV;
  constexpr int LW = NV * VEC;
  constexpr int TW = 64 * VEC, TH = 4 * PPT;
  constexpr int LH = TH + 2 * RY;
  constexpr int NW = 1 + 2 * HV;
  constexpr bool IS_F = std::is_same<ACC, float>::value;
  __shared__ __align__(16) TI tile[LH * LW];

  const Geom& g = a.g;
  const int tx0 = blockIdx.x * TW;
  const int ty0 = blockIdx.y * TH;
  const TI* in = static_cast<const TI*>(g.in) + (long long)blockIdx.z * g.in_plane;
  TO* out = static_cast<TO*>(g.out) + (long long)blockIdx.z * g.out_plane;
  const int iy0 = g.in_oy + ty0 - RY;
  stage_tile<TI, VEC, NV, LH>(tile, in, g, g.in_ox + tx0 - HV * VEC, iy0, a.in_vec_ok);
  __syncthreads();

  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const bool cmode = g.boundary == HIPACC_BOUNDARY_CONSTANT;
  const ACC cval = IS_F ? (ACC)g.const_f : (ACC)g.const_i;
  ACC acc[PPT][VEC];
#pragma unroll
  for (int o = 0; o < PPT; ++o)
#pragma unroll
    for (int v = 0; v < VEC; ++v) acc[o][v] = ACC(0);

  const TI* base = tile + (wave * PPT) * LW + lane * VEC;
#pragma unroll
  for (int ir = 0; ir < PPT + 2 * RY; ++ir) {
    ACC win[NW * VEC];
    load_window<TI, ACC, VEC, NW>(win, base + ir * LW);
    ACC h[VEC];
    if constexpr (BOX) {
      // All-ones integer row: running window sums, exact modulo 2^32.
      ACC s0 = ACC(0);
#pragma unroll
      for (int dx = 0; dx < KX; ++dx) s0 = (ACC)((uint32_t)s0 + (uint32_t)win[HV * VEC + dx - RX]);
      h[0] = s0;
#pragma unroll
      for (int v = 1; v < VEC; ++v)
        h[v] = (ACC)((uint32_t)h[v - 1] + (uint32_t)win[HV * VEC + v + RX] -
                     (uint32_t)win[HV * VEC + v - 1 - RX]);
    } else {
#pragma unroll
      for (int v = 0; v < VEC; ++v) {
        ACC s0 = ACC(0);
#pragma unroll
        for (int dx = 0; dx < KX; ++dx) {
          const ACC w = IS_F ? (ACC)a.rowf[dx] : (ACC)a.rowi[dx];
          s0 = madd<FMA>(w, win[HV * VEC + v + dx - RX], s0);
        }
        h[v] = s0;
      }
    }
    if (cmode) {
      const int yy = iy0 + wave * PPT + ir;
      if (yy < g.lo_y || yy >= g.hi_y) {
#pragma unroll
        for (int v = 0; v < VEC; ++v) h[v] = cval;
      }
    }
#pragma unroll
    for (int o = 0; o < PPT; ++o) {
      const int dy = ir - o;
      if (dy < 0 || dy >= KY) continue;
      if constexpr (BOX) {
#pragma unroll
        for (int v = 0; v < VEC; ++v) acc[o][v] = (ACC)((uint32_t)acc[o][v] + (uint32_t)h[v]);
      } else {
        const ACC w = IS_F ? (ACC)a.colf[dy] : (ACC)a.coli[dy];
#pragma unroll
        for (int v = 0; v < VEC; ++v) acc[o][v] = madd<FMA>(w, h[v], acc[o][v]);
      }
    }
  }

  const int ox = tx0 + lane * VEC;
#pragma unroll
  for (int o = 0; o < PPT; ++o)
    store_row<TO, ACC, VEC>(out, g, a.out_vec_ok, ox, ty0 + wave * PPT + o, acc[o]);
}

template <typename TI, typename TO, typename ACC, int KX, int KY, int PPT, bool FMA, bool BOX>
static int launch_sep(const SepArgs& a, hipStream_t s) {
  constexpr int VEC = 16 / (int)sizeof(TI);
  dim3 grid(ceil_div(a.g.is_w, 64 * VEC), ceil_div(a.g.is_h, 4 * PPT), a.g.batch);
  hipLaunchKernelGGL((separable_kernel<TI, TO, ACC, KX, KY, PPT, FMA, BOX>), grid, dim3(256), 0, s,
                     a);
  return hip_status(hipGetLastError(), "separable_kernel launch");
}

#define HH_SEPF(TI, TO, DTI, DTO, K, P)                                                    \
  if (a.g.in_dtype == DTI && a.g.out_dtype == DTO && a.kx == K && a.ky == K && ppt == P) {  \
    if (fma) return launch_sep<TI, TO, float, K, K, P, true, false>(a, s);                \
    return launch_sep<TI, TO, float, K, K, P, false, false>(a, s);                        \
  }
#define HH_SEPI(TO, DTO, K, P)                                                                \
  if (a.g.in_dtype == HIPACC_U8 && a.g.out_dtype == DTO && a.kx == K && a.ky == K && ppt == P) { \
    if (a.box) return launch_sep<uint8_t, TO, int32_t, K, K, P, false, true>(a, s);           \
    return launch_sep<uint8_t, TO, int32_t, K, K, P, false, false>(a, s);                     \
  }
#define HH_SEPF_K(K)                                                          \
  HH_SEPF(float, float, HIPACC_F32, HIPACC_F32, K, 8)                         \
  HH_SEPF(float, float, HIPACC_F32, HIPACC_F32, K, 4)                         \
  HH_SEPF(uint8_t, uint8_t, HIPACC_U8, HIPACC_U8, K, 4)                       \
  HH_SEPF(uint8_t, uint8_t, HIPACC_U8, HIPACC_U8, K, 2)
#define HH_SEPI_K(K)                                                          \
  HH_SEPI(uint32_t, HIPACC_U32, K, 4) HH_SEPI(uint32_t, HIPACC_U32, K, 2)     \
  HH_SEPI(int32_t, HIPACC_I32, K, 4) HH_SEPI(int16_t, HIPACC_I16, K, 4)     \
  HH_SEPI(uint8_t, HIPACC_U8, K, 4)

static int dispatch_sep(const SepArgs& a, int acc_is_float, int fma, int ppt, hipStream_t s) {
  if (acc_is_float) {
    if (ppt == 0) ppt = a.g.in_dtype == HIPACC_F32 ? 8 : 4;
    HH_SEPF_K(3)
    HH_SEPF_K(5)
    HH_SEPF_K(7)
    HH_SEPF_K(9)
    HH_SEPF_K(13)
  } else {
    if (ppt == 0) ppt = 4;
    HH_SEPI_K(3)
    HH_SEPI_K(5)
    HH_SEPI_K(7)
  }
  return HIPACC_ERR_UNSUPPORTED;
}

// Generic two-pass form for shapes without an instantiation: the row pass
// writes a float/int intermediate through the generic local kernel's rules.
template <typename ACC>
__global__ __launch_bounds__(256) void separable_generic_kernel(const SepArgs a) {
  const Geom& g = a.g;
  const int x = blockIdx.x * 64 + threadIdx.x, y = blockIdx.y * 4 + threadIdx.y;
  if (x >= g.is_w || y >= g.is_h) return;
  constexpr bool IS_F = std::is_same<ACC, float>::value;
  const char* in = static_cast<const char*>(g.in) +
                   (long long)blockIdx.z * g.in_plane * dtype_bytes(g.in_dtype);
  char* out = static_cast<char*>(g.out) +
              (long long)blockIdx.z * g.out_plane * dtype_bytes(g.out_dtype);
  const int rx = a.kx / 2, ry = a.ky / 2;
  const bool cmode = g.boundary == HIPACC_BOUNDARY_CONSTANT;
  const ACC cval = IS_F ? (ACC)g.const_f : (ACC)g.const_i;
  const int ix = g.in_ox + x, iy = g.in_oy + y;
  ACC acc = ACC(0);
  for (int dy = 0; dy < a.ky; ++dy) {
    int yy = iy + dy - ry;
    ACC h = ACC(0);
    if (cmode && (yy < g.lo_y || yy >= g.hi_y)) {
      h = cval;
    } else {
      if (!cmode) yy = bh_map(g.boundary, yy, g.lo_y, g.hi_y, g.in_imgh);
      for (int dx = 0; dx < a.kx; ++dx) {
        int xx = ix + dx - rx;
        ACC v;
        if (cmode && (xx < g.lo_x || xx >= g.hi_x)) {
          v = cval;
        } else {
          if (!cmode) xx = bh_map(g.boundary, xx, g.lo_x, g.hi_x, g.in_imgw);
          const long long idx = (long long)yy * g.in_stride + xx;
          if constexpr (IS_F) v = load_f(in, g.in_dtype, idx); else v = load_i(in, g.in_dtype, idx);
        }
        if constexpr (IS_F) {
          h = a.box ? madd<true>(a.rowf[dx], v, h) : madd<false>(a.rowf[dx], v, h);
        } else {
          h = madd<false>(a.rowi[dx], v, h);
        }
      }
    }
    if constexpr (IS_F) {
      acc = a.box ? madd<true>(a.colf[dy], h, acc) : madd<false>(a.colf[dy], h, acc);
    } else {
      acc = madd<false>(a.coli[dy], h, acc);
    }
  }
  const long long oidx = (long long)(g.out_oy + y) * g.out_stride + g.out_ox + x;
  if constexpr (IS_F) store_f(out, g.out_dtype, oidx, acc, g.cvt, g.post_div);
  else store_i(out, g.out_dtype, oidx, acc, g.cvt, g.post_div);
}

static int launch_sep_generic(const SepArgs& a, int acc_is_float, int fp_contract, hipStream_t s) {
  // The generic kernel reuses `box` as the float contraction flag.
  SepArgs b = a;
  if (acc_is_float) b.box = fp_contract;
  dim3 grid(ceil_div(a.g.is_w, 64), ceil_div(a.g.is_h, 4), a.g.batch);
  if (acc_is_float)
    hipLaunchKernelGGL(separable_generic_kernel<float>, grid, dim3(64, 4), 0, s, b);
  else
    hipLaunchKernelGGL(separable_generic_kernel<int32_t>, grid, dim3(64, 4), 0, s, b);
  return hip_status(hipGetLastError(), "separable_generic_kernel launch");
}

static int launch_sep_variant(const SepArgs& a, int acc_is_float, int fma, hipStream_t s,
                              int variant, int param) {
  switch (variant) {
    case kVarStream: return dispatch_stream_sep(a, acc_is_float, fma, param, 0, s);
    case kVarStreamWide: return dispatch_stream_sep(a, acc_is_float, fma, param, 1, s);
    case kVarTile: return dispatch_sep(a, acc_is_float, fma, param, s);
    case kVarGeneric: return launch_sep_generic(a, acc_is_float, fma, s);
    case kVarRing: return dispatch_ring_sep(a, acc_is_float, fma, param, s);
  }
  return HIPACC_ERR_UNSUPPORTED;
}

// variant kVarAuto: the tuned choice if any, else ring -> stream -> tile ->
// generic.
int launch_separable(const SepArgs& a, int acc_is_float, int fp_contract, hipStream_t s,
                     int variant, int param) {
  if (variant != kVarAuto) return launch_sep_variant(a, acc_is_float, fp_contract, s, variant, param);
  TuneKey k{a.box ? kTuneBox : kTuneSeparable, a.kx, a.ky, a.g.in_dtype, a.g.out_dtype,
            acc_is_float, a.g.boundary, fp_contract, 0,
            tune_size_class(a.g.is_w, a.g.is_h, a.g.batch)};
  TuneVal v;
  if (tune_lookup(k, v)) {
    const int st = launch_sep_variant(a, acc_is_float, fp_contract, s, v.variant, v.param);
    if (st != HIPACC_ERR_UNSUPPORTED) return st;
  }
  int st = launch_sep_variant(a, acc_is_float, fp_contract, s, kVarRing, 0);
  if (st == HIPACC_ERR_UNSUPPORTED) st = launch_sep_variant(a, acc_is_float, fp_contract, s, kVarStream, 0);
  if (st == HIPACC_ERR_UNSUPPORTED) st = launch_sep_variant(a, acc_is_float, fp_contract, s, kVarTile, 0);
  if (st == HIPACC_ERR_UNSUPPORTED) st = launch_sep_generic(a, acc_is_float, fp_contract, s);
  return st;
}

}  // namespace hh
