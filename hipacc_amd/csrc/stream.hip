// stream.hip -- instantiations and launch geometry of the streaming local
// operators (hh_stream.hpp).
#include "hh_stream.hpp"

namespace hh {

// Alignment the streaming kernels need: 4-pixel loads and stores.
static void stream_align(const Geom& g, int& in_ok, int& out_ok) {
  const int ebi = dtype_bytes(g.in_dtype), ebo = dtype_bytes(g.out_dtype);
  const uintptr_t pi = reinterpret_cast<uintptr_t>(g.in), po = reinterpret_cast<uintptr_t>(g.out);
  in_ok = (pi % (4 * ebi)) == 0 && (g.in_stride % 4) == 0 && (g.in_ox % 4) == 0 &&
          (g.batch <= 1 || (g.in_plane % 4) == 0);
  out_ok = (po % (4 * ebo)) == 0 && (g.out_stride % 4) == 0 && (g.out_ox % 4) == 0 &&
           (g.batch <= 1 || (g.out_plane % 4) == 0);
}

static StreamGeom stream_geom(const Geom& g, int rx, int seg_rows) {
  StreamGeom sg;
  const int hl = (rx + kStreamVec - 1) / kStreamVec;
  sg.strip_w = (64 - 2 * hl) * kStreamVec;
  sg.strips = ceil_div(g.is_w, sg.strip_w);
  sg.seg_rows = seg_rows > 0 ? seg_rows : 32;
  sg.segs = ceil_div(g.is_h, sg.seg_rows);
  sg.total_waves = sg.strips * sg.segs * g.batch;
  sg.xcd_perm = 0;
  return sg;
}

static dim3 stream_grid(StreamGeom& sg) {
  int blocks = ceil_div(sg.total_waves, 4);
  if (blocks >= 64) {
    blocks = (blocks + 7) / 8 * 8;
    sg.xcd_perm = 1;
  }
  return dim3(blocks);
}

template <typename TI, typename TO, typename ACC, int KX, int KY, int PF, bool FMA, int KIND,
          bool WDEV, typename ArgsT>
static int launch_stream(const ArgsT& a0, int seg_rows, hipStream_t s) {
  ArgsT a = a0;
  stream_align(a.g, a.in_vec_ok, a.out_vec_ok);
  StreamGeom sg = stream_geom(a.g, KX / 2, seg_rows);
  const dim3 grid = stream_grid(sg);
  hipLaunchKernelGGL((stream_kernel<TI, TO, ACC, KX, KY, PF, FMA, KIND, WDEV, ArgsT>), grid,
                     dim3(256), 0, s, a, sg);
  return hip_status(hipGetLastError(), "stream_kernel launch");
}

// Prefetch depth: at least ~8 rows, a multiple of KY.
#define HH_PF(K) ((K) >= 7 ? (K) : ((K) == 5 ? 10 : 9))

int dispatch_stream_local(const LocalArgs& a, int acc_is_float, int seg, hipStream_t s) {
  if (a.reduce != HIPACC_REDUCE_SUM || a.mask_w != a.mask_h) return HIPACC_ERR_UNSUPPORTED;
  const Geom& g = a.g;
  const bool wd = a.wdev != nullptr;
#define HH_S2F(K)                                                                              \
  if (a.mask_w == K) {                                                                         \
    if (a.fma) return wd ? launch_stream<float, float, float, K, K, HH_PF(K), true, kStream2D, true>(a, seg, s) \
                         : launch_stream<float, float, float, K, K, HH_PF(K), true, kStream2D, false>(a, seg, s); \
    return wd ? launch_stream<float, float, float, K, K, HH_PF(K), false, kStream2D, true>(a, seg, s)         \
              : launch_stream<float, float, float, K, K, HH_PF(K), false, kStream2D, false>(a, seg, s);       \
  }
#define HH_S2U(K)                                                                              \
  if (a.mask_w == K) {                                                                         \
    if (a.fma) return launch_stream<uint8_t, uint8_t, float, K, K, HH_PF(K), true, kStream2D, false>(a, seg, s); \
    return launch_stream<uint8_t, uint8_t, float, K, K, HH_PF(K), false, kStream2D, false>(a, seg, s);         \
  }
#define HH_S2I(TO, DT, K)                                                                      \
  if (g.out_dtype == DT && a.mask_w == K)                                                      \
    return launch_stream<uint8_t, TO, int32_t, K, K, HH_PF(K), false, kStream2D, false>(a, seg, s);
  if (acc_is_float) {
    if (g.in_dtype == HIPACC_F32 && g.out_dtype == HIPACC_F32) {
      HH_S2F(3) HH_S2F(5) HH_S2F(7) HH_S2F(9) HH_S2F(13)
    } else if (g.in_dtype == HIPACC_U8 && g.out_dtype == HIPACC_U8 && !wd) {
      HH_S2U(3) HH_S2U(5) HH_S2U(7) HH_S2U(9)
    }
  } else if (g.in_dtype == HIPACC_U8 && !wd) {
    HH_S2I(int16_t, HIPACC_I16, 3) HH_S2I(int16_t, HIPACC_I16, 5) HH_S2I(int16_t, HIPACC_I16, 7)
    HH_S2I(uint32_t, HIPACC_U32, 3) HH_S2I(uint32_t, HIPACC_U32, 5) HH_S2I(uint32_t, HIPACC_U32, 7)
    HH_S2I(int32_t, HIPACC_I32, 3) HH_S2I(int32_t, HIPACC_I32, 5) HH_S2I(int32_t, HIPACC_I32, 7)
  }
  return HIPACC_ERR_UNSUPPORTED;
}

int dispatch_stream_sep(const SepArgs& a, int acc_is_float, int fma, int seg, hipStream_t s) {
  if (a.kx != a.ky) return HIPACC_ERR_UNSUPPORTED;
  const Geom& g = a.g;
#define HH_SSF(TI, DT, K)                                                                       \
  if (g.in_dtype == DT && g.out_dtype == DT && a.kx == K) {                                     \
    if (fma) return launch_stream<TI, TI, float, K, K, HH_PF(K), true, kStreamSep, false>(a, seg, s); \
    return launch_stream<TI, TI, float, K, K, HH_PF(K), false, kStreamSep, false>(a, seg, s);        \
  }
#define HH_SSB(TO, DT, K)                                                                       \
  if (g.in_dtype == HIPACC_U8 && g.out_dtype == DT && a.kx == K)                                \
    return launch_stream<uint8_t, TO, int32_t, K, K, HH_PF(K), false, kStreamBox, false>(a, seg, s);
  if (acc_is_float) {
    HH_SSF(float, HIPACC_F32, 3) HH_SSF(float, HIPACC_F32, 5) HH_SSF(float, HIPACC_F32, 7)
    HH_SSF(float, HIPACC_F32, 9) HH_SSF(float, HIPACC_F32, 13)
    HH_SSF(uint8_t, HIPACC_U8, 3) HH_SSF(uint8_t, HIPACC_U8, 5) HH_SSF(uint8_t, HIPACC_U8, 7)
    HH_SSF(uint8_t, HIPACC_U8, 9) HH_SSF(uint8_t, HIPACC_U8, 13)
  } else if (a.box) {
    HH_SSB(uint32_t, HIPACC_U32, 3) HH_SSB(uint32_t, HIPACC_U32, 5) HH_SSB(uint32_t, HIPACC_U32, 7)
    HH_SSB(int32_t, HIPACC_I32, 3) HH_SSB(int32_t, HIPACC_I32, 5) HH_SSB(int32_t, HIPACC_I32, 7)
    HH_SSB(int16_t, HIPACC_I16, 3) HH_SSB(int16_t, HIPACC_I16, 5) HH_SSB(int16_t, HIPACC_I16, 7)
    HH_SSB(uint8_t, HIPACC_U8, 3) HH_SSB(uint8_t, HIPACC_U8, 5) HH_SSB(uint8_t, HIPACC_U8, 7)
  }
  return HIPACC_ERR_UNSUPPORTED;
}

}  // namespace hh
