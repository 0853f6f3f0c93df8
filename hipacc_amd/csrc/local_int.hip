// local_int.hip -- fast integer weighted-SUM instantiations: uchar image with
// an int accumulator into short / uint / int images (Sobel, box sums).
#include "hh_local_fast.hpp"

namespace hh {

#define HH_INT_CASE(TO, DT, M, P)                                                  \
  if (a.g.out_dtype == DT && a.mask_w == M && a.mask_h == M && ppt == P)           \
    return launch_sum<uint8_t, TO, int32_t, M, M, P, false, false>(a, s);
#define HH_INT_SIZE(M)                                                                      \
  HH_INT_CASE(int16_t, HIPACC_I16, M, 2) HH_INT_CASE(int16_t, HIPACC_I16, M, 4)          \
  HH_INT_CASE(uint32_t, HIPACC_U32, M, 2) HH_INT_CASE(uint32_t, HIPACC_U32, M, 4)        \
  HH_INT_CASE(int32_t, HIPACC_I32, M, 2) HH_INT_CASE(int32_t, HIPACC_I32, M, 4)

int dispatch_sum_int(const LocalArgs& a, int ppt, hipStream_t s) {
  if (a.g.in_dtype != HIPACC_U8 || a.wdev) return HIPACC_ERR_UNSUPPORTED;
  if (ppt == 0) ppt = 4;
  HH_INT_SIZE(3)
  HH_INT_SIZE(5)
  HH_INT_SIZE(7)
  return HIPACC_ERR_UNSUPPORTED;
}

}  // namespace hh
