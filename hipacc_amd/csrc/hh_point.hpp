// hh_point.hpp -- the per-pixel arithmetic of the point operators, shared by
// point_kernel (point.hip) and the point stages of fused chains (fused.hip).
#pragma once

#include "hh_common.hpp"

namespace hh {

template <typename A> __device__ __forceinline__ A point_eval(int op, A a, A b, float alpha, float beta);
template <> __device__ __forceinline__ float point_eval<float>(int op, float a, float b, float alpha,
                                                               float beta) {
  switch (op) {
    case HIPACC_POINT_AFFINE: return alpha * a + beta;
    case HIPACC_POINT_ADD: return a + b;
    case HIPACC_POINT_SUB: return a - b;
    case HIPACC_POINT_MUL: return a * b;
    case HIPACC_POINT_ABSDIFF: return fabsf(a - b);
    case HIPACC_POINT_MIN: return a < b ? a : b;
    case HIPACC_POINT_MAX: return a > b ? a : b;
    case HIPACC_POINT_MAGNITUDE: {
      float r = sqrtf(a * a + b * b);
      r = r < 255.0f ? r : 255.0f;
      return r > 0.0f ? r : 0.0f;
    }
    default: return a;
  }
}
template <> __device__ __forceinline__ int32_t point_eval<int32_t>(int op, int32_t a, int32_t b,
                                                                   float alpha, float beta) {
  switch (op) {
    case HIPACC_POINT_ADD: return (int32_t)((uint32_t)a + (uint32_t)b);
    case HIPACC_POINT_SUB: return (int32_t)((uint32_t)a - (uint32_t)b);
    case HIPACC_POINT_MUL: return (int32_t)((uint32_t)a * (uint32_t)b);
    case HIPACC_POINT_ABSDIFF: return a > b ? a - b : b - a;
    case HIPACC_POINT_MIN: return a < b ? a : b;
    case HIPACC_POINT_MAX: return a > b ? a : b;
    default: return a;
  }
}

// One output pixel of a point operator, converted to TO.  FLOAT_MATH (a
// float input) evaluates in float; integer inputs use int arithmetic except
// AFFINE (float) and MAGNITUDE (int division by the norm alpha, int sum of
// squares, float sqrt and clamp to [0, 255]).  af/bf and ai/bi are the
// inputs converted to float and to int32.
template <typename TO, bool FLOAT_MATH>
__device__ __forceinline__ TO point_pixel(int op, float af, float bf, int32_t ai, int32_t bi,
                                          float alpha, float beta, int cvt) {
  if constexpr (FLOAT_MATH) {
    return convert_f<TO>(point_eval<float>(op, af, bf, alpha, beta), cvt, 0.0f);
  } else {
    if (op == HIPACC_POINT_AFFINE) return convert_f<TO>(alpha * af + beta, cvt, 0.0f);
    if (op == HIPACC_POINT_MAGNITUDE) {
      const int32_t nrm = (int32_t)alpha;
      const int32_t q0 = nrm ? ai / nrm : ai;
      const int32_t q1 = nrm ? bi / nrm : bi;
      float f = sqrtf((float)(q0 * q0 + q1 * q1));
      f = f < 255.0f ? f : 255.0f;
      f = f > 0.0f ? f : 0.0f;
      return convert_f<TO>(f, cvt, 0.0f);
    }
    return convert_i<TO>(point_eval<int32_t>(op, ai, bi, 0.f, 0.f), cvt, 0.0f);
  }
}

}  // namespace hh
