// hh_fused.hpp -- argument block of a fused operator chain (hipacc_hip_fused),
// resolved on the host (capi.hip) and read by fused_kernel (fused.hip).
#pragma once

#include "hh_local.hpp"

namespace hh {

// A second input of a point stage; its window has the chain's size.
struct FusedImg {
  const void* data;        // NULL for a one-input op
  long long plane;         // elements between batch images
  int stride, ox, oy, dtype;
};
struct FusedPoint {
  int op, cvt, dtype;      // dtype: the pixel type the stage writes
  float alpha, beta;
  FusedImg in1;
};
struct FusedLocal {
  int mask_w, mask_h, reduce, use_weights, fma, boundary, cvt;
  int in_dtype, out_dtype, acc_f;
  float post_div;
  float const_f;           // CONSTANT, in the pixel type of the local's input
  int const_i;
  uint8_t domain[kMaxTaps];
  Weights w;
};
struct FusedArgs {
  const void* in;
  void* out;
  long long in_plane, out_plane;
  int in_stride, out_stride, in_dtype, out_dtype;
  int in_imgw, in_imgh, in_ox, in_oy, out_ox, out_oy;
  int w, h, batch;
  int num_locals;
  // Point stages in chain order; seg[s] = [first, end) of the points before
  // the first local (s = 0), between the locals (1) and after the last (2).
  int seg[3][2];
  FusedPoint pt[HIPACC_MAX_FUSED_STAGES];
  FusedLocal loc[HIPACC_MAX_FUSED_LOCALS];
};
static_assert(sizeof(FusedArgs) <= 4096, "fused chain arguments exceed 4 KiB of kernarg");

int launch_fused(const FusedArgs& a, hipStream_t s);

}  // namespace hh
