// stream_2d.hip -- instantiations of the streaming 2D-mask operators
// (hh_stream.hpp).  `wide` selects 8 pixels per lane instead of 4.
#include "hh_stream.hpp"

namespace hh {

int dispatch_stream_local(const LocalArgs& a, int acc_is_float, int seg, int wide, hipStream_t s) {
  if (a.reduce != HIPACC_REDUCE_SUM || a.mask_w != a.mask_h) return HIPACC_ERR_UNSUPPORTED;
  const Geom& g = a.g;
  const bool wd = a.wdev != nullptr;
#define HH_S2F(V, K)                                                                             \
  if (a.mask_w == K) {                                                                           \
    if (a.fma) return wd ? launch_stream<float, float, float, V, K, K, HH_PF(K), true, kStream2D, true>(a, seg, s) \
                         : launch_stream<float, float, float, V, K, K, HH_PF(K), true, kStream2D, false>(a, seg, s); \
    return wd ? launch_stream<float, float, float, V, K, K, HH_PF(K), false, kStream2D, true>(a, seg, s)         \
              : launch_stream<float, float, float, V, K, K, HH_PF(K), false, kStream2D, false>(a, seg, s);       \
  }
#define HH_S2U(V, K)                                                                             \
  if (a.mask_w == K) {                                                                           \
    if (a.fma) return launch_stream<uint8_t, uint8_t, float, V, K, K, HH_PF(K), true, kStream2D, false>(a, seg, s); \
    return launch_stream<uint8_t, uint8_t, float, V, K, K, HH_PF(K), false, kStream2D, false>(a, seg, s);         \
  }
#define HH_S2I(TO, DT, V, K)                                                                     \
  if (g.out_dtype == DT && a.mask_w == K)                                                        \
    return launch_stream<uint8_t, TO, int32_t, V, K, K, HH_PF(K), false, kStream2D, false>(a, seg, s);
  if (acc_is_float) {
    if (g.in_dtype == HIPACC_F32 && g.out_dtype == HIPACC_F32) {
      if (wide) {
        if (!wd) { HH_S2F(8, 3) HH_S2F(8, 5) HH_S2F(8, 7) }
      } else {
        HH_S2F(4, 3) HH_S2F(4, 5) HH_S2F(4, 7) HH_S2F(4, 9) HH_S2F(4, 13)
      }
    } else if (g.in_dtype == HIPACC_U8 && g.out_dtype == HIPACC_U8 && !wd) {
      if (wide) { HH_S2U(8, 3) HH_S2U(8, 5) }
      else { HH_S2U(4, 3) HH_S2U(4, 5) HH_S2U(4, 7) HH_S2U(4, 9) }
    }
  } else if (g.in_dtype == HIPACC_U8 && !wd) {
    if (wide) {
      HH_S2I(int16_t, HIPACC_I16, 8, 3) HH_S2I(uint32_t, HIPACC_U32, 8, 3) HH_S2I(int32_t, HIPACC_I32, 8, 3)
    } else {
      HH_S2I(int16_t, HIPACC_I16, 4, 3) HH_S2I(int16_t, HIPACC_I16, 4, 5) HH_S2I(int16_t, HIPACC_I16, 4, 7)
      HH_S2I(uint32_t, HIPACC_U32, 4, 3) HH_S2I(uint32_t, HIPACC_U32, 4, 5) HH_S2I(uint32_t, HIPACC_U32, 4, 7)
      HH_S2I(int32_t, HIPACC_I32, 4, 3) HH_S2I(int32_t, HIPACC_I32, 4, 5) HH_S2I(int32_t, HIPACC_I32, 4, 7)
    }
  }
  return HIPACC_ERR_UNSUPPORTED;
}

}  // namespace hh
