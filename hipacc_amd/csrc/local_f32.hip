// local_f32.hip -- fast weighted-SUM instantiations, float image -> float image.
#include "hh_local_fast.hpp"

namespace hh {

#define HH_F32_CASE(M, P, F, W)                                                         \
  if (a.mask_w == M && a.mask_h == M && ppt == P && a.fma == F && (a.wdev != nullptr) == W) \
    return launch_sum<float, float, float, M, M, P, F, W>(a, s);
#define HH_F32_SIZE(M)     \
  HH_F32_CASE(M, 4, 0, 0)  \
  HH_F32_CASE(M, 4, 1, 0)  \
  HH_F32_CASE(M, 8, 0, 0)  \
  HH_F32_CASE(M, 8, 1, 0)  \
  HH_F32_CASE(M, 4, 0, 1)  \
  HH_F32_CASE(M, 4, 1, 1)  \
  HH_F32_CASE(M, 8, 0, 1)  \
  HH_F32_CASE(M, 8, 1, 1)

int dispatch_sum_f32(const LocalArgs& a, int ppt, hipStream_t s) {
  if (a.g.in_dtype != HIPACC_F32 || a.g.out_dtype != HIPACC_F32) return HIPACC_ERR_UNSUPPORTED;
  if (ppt == 0) ppt = 8;
  HH_F32_SIZE(3)
  HH_F32_SIZE(5)
  HH_F32_SIZE(7)
  HH_F32_SIZE(9)
  HH_F32_SIZE(13)
  return HIPACC_ERR_UNSUPPORTED;
}

}  // namespace hh
