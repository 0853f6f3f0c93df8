// local_u8.hip -- fast weighted-SUM instantiations, uchar image -> uchar image
// with a float accumulator (Gaussian-style masks).
#include "hh_local_fast.hpp"

namespace hh {

#define HH_U8_CASE(M, P, F)                                                                \
  if (a.mask_w == M && a.mask_h == M && ppt == P && a.fma == F)                            \
    return launch_sum<uint8_t, uint8_t, float, M, M, P, F, false>(a, s);
#define HH_U8_SIZE(M) HH_U8_CASE(M, 2, 0) HH_U8_CASE(M, 2, 1) HH_U8_CASE(M, 4, 0) HH_U8_CASE(M, 4, 1)

int dispatch_sum_u8(const LocalArgs& a, int ppt, hipStream_t s) {
  if (a.g.in_dtype != HIPACC_U8 || a.g.out_dtype != HIPACC_U8 || a.wdev) return HIPACC_ERR_UNSUPPORTED;
  if (ppt == 0) ppt = 4;
  HH_U8_SIZE(3)
  HH_U8_SIZE(5)
  HH_U8_SIZE(7)
  HH_U8_SIZE(9)
  HH_U8_SIZE(13)
  return HIPACC_ERR_UNSUPPORTED;
}

}  // namespace hh
