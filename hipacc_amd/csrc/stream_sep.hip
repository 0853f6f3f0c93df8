// stream_sep.hip -- instantiations of the streaming separable and box-sum
// operators (hh_stream.hpp).  `wide` selects 8 pixels per lane instead of 4.
#include "hh_stream.hpp"

namespace hh {

int dispatch_stream_sep(const SepArgs& a, int acc_is_float, int fma, int seg, int wide, hipStream_t s) {
  if (a.kx != a.ky) return HIPACC_ERR_UNSUPPORTED;
  const Geom& g = a.g;
#define HH_SSF(TI, DT, V, K)                                                                      \
  if (g.in_dtype == DT && g.out_dtype == DT && a.kx == K) {                                       \
    if (fma) return launch_stream<TI, TI, float, V, K, K, HH_PF(K), true, kStreamSep, false>(a, seg, s); \
    return launch_stream<TI, TI, float, V, K, K, HH_PF(K), false, kStreamSep, false>(a, seg, s);        \
  }
#define HH_SSB(TO, DT, V, K)                                                                      \
  if (g.in_dtype == HIPACC_U8 && g.out_dtype == DT && a.kx == K)                                  \
    return launch_stream<uint8_t, TO, int32_t, V, K, K, HH_PF(K), false, kStreamBox, false>(a, seg, s);
  if (acc_is_float) {
    if (wide) {
      HH_SSF(float, HIPACC_F32, 8, 5) HH_SSF(float, HIPACC_F32, 8, 13)
      HH_SSF(uint8_t, HIPACC_U8, 8, 5) HH_SSF(uint8_t, HIPACC_U8, 8, 13)
    } else {
      HH_SSF(float, HIPACC_F32, 4, 3) HH_SSF(float, HIPACC_F32, 4, 5) HH_SSF(float, HIPACC_F32, 4, 7)
      HH_SSF(float, HIPACC_F32, 4, 9) HH_SSF(float, HIPACC_F32, 4, 13)
      HH_SSF(uint8_t, HIPACC_U8, 4, 3) HH_SSF(uint8_t, HIPACC_U8, 4, 5) HH_SSF(uint8_t, HIPACC_U8, 4, 7)
      HH_SSF(uint8_t, HIPACC_U8, 4, 9) HH_SSF(uint8_t, HIPACC_U8, 4, 13)
    }
  } else if (a.box) {
    if (wide) {
      HH_SSB(uint32_t, HIPACC_U32, 8, 7) HH_SSB(uint32_t, HIPACC_U32, 8, 3)
    } else {
      HH_SSB(uint32_t, HIPACC_U32, 4, 3) HH_SSB(uint32_t, HIPACC_U32, 4, 5) HH_SSB(uint32_t, HIPACC_U32, 4, 7)
      HH_SSB(int32_t, HIPACC_I32, 4, 3) HH_SSB(int32_t, HIPACC_I32, 4, 5) HH_SSB(int32_t, HIPACC_I32, 4, 7)
      HH_SSB(int16_t, HIPACC_I16, 4, 3) HH_SSB(int16_t, HIPACC_I16, 4, 5) HH_SSB(int16_t, HIPACC_I16, 4, 7)
      HH_SSB(uint8_t, HIPACC_U8, 4, 3) HH_SSB(uint8_t, HIPACC_U8, 4, 5) HH_SSB(uint8_t, HIPACC_U8, 4, 7)
    }
  }
  return HIPACC_ERR_UNSUPPORTED;
}

}  // namespace hh
