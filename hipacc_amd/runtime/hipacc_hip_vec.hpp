// hipacc_hip_vec.hpp -- the DSL's 4-wide pixel types on HIP's vector types.
//
// The DSL declares char4 ... double4 with componentwise arithmetic,
// make_<type>4, the convert_<type>4 conversions and componentwise math
// (dsl/types.hpp:38-100,486-516, dsl/math_functions.hpp:44-404).  HIP's
// vector types already carry the arithmetic and the make_ functions; this
// header adds the rest for host and device code alike (the programs' CPU
// reference code uses them too):
//   - pixel codes for images of uchar4 / short4 / int4 / float4 pixels
//     (HIPACC_U8X4 ... HIPACC_F32X4, four interleaved channels);
//   - convert_<type>4(v): a C cast of each component;
//   - hipacc::math::min / max (vector-vector, vector-scalar, scalar-vector)
//     and abs, componentwise;
//   - hipacc::math::<f>f(float4) for the single-precision math functions
//     (expf, sqrtf, ...), componentwise.
#ifndef HIPACC_HIP_VEC_HPP
#define HIPACC_HIP_VEC_HPP

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <type_traits>

#include "hipacc_hip.hpp"

template <> struct hipacc_pixel_code<uchar4> { static constexpr int value = HIPACC_U8X4; };
template <> struct hipacc_pixel_code<short4> { static constexpr int value = HIPACC_I16X4; };
template <> struct hipacc_pixel_code<int4> { static constexpr int value = HIPACC_I32X4; };
template <> struct hipacc_pixel_code<float4> { static constexpr int value = HIPACC_F32X4; };

#define HIPACC_VEC_FN __host__ __device__ inline

#define HIPACC_VEC_CONVERT(NAME, E)                                                       \
  template <typename V>                                                                   \
  HIPACC_VEC_FN NAME convert_##NAME(const V& v) {                                         \
    return make_##NAME((E)v.x, (E)v.y, (E)v.z, (E)v.w);                                   \
  }
HIPACC_VEC_CONVERT(char4, signed char)
HIPACC_VEC_CONVERT(uchar4, unsigned char)
HIPACC_VEC_CONVERT(short4, short)
HIPACC_VEC_CONVERT(ushort4, unsigned short)
HIPACC_VEC_CONVERT(int4, int)
HIPACC_VEC_CONVERT(uint4, unsigned int)
HIPACC_VEC_CONVERT(long4, long)
HIPACC_VEC_CONVERT(ulong4, unsigned long)
HIPACC_VEC_CONVERT(float4, float)
HIPACC_VEC_CONVERT(double4, double)
#undef HIPACC_VEC_CONVERT

namespace hipacc {
namespace math {

template <typename E> using vec4 = HIP_vector_type<E, 4>;

template <typename E> HIPACC_VEC_FN E vmin1(E a, E b) { return a < b ? a : b; }
template <typename E> HIPACC_VEC_FN E vmax1(E a, E b) { return a > b ? a : b; }

template <typename E> HIPACC_VEC_FN vec4<E> min(vec4<E> a, vec4<E> b) {
  return vec4<E>(vmin1(a.x, b.x), vmin1(a.y, b.y), vmin1(a.z, b.z), vmin1(a.w, b.w));
}
template <typename E> HIPACC_VEC_FN vec4<E> max(vec4<E> a, vec4<E> b) {
  return vec4<E>(vmax1(a.x, b.x), vmax1(a.y, b.y), vmax1(a.z, b.z), vmax1(a.w, b.w));
}
// With a scalar: the scalar is converted to the component type, as in the
// DSL's min(NEW_TYPE, BASIC_TYPE).
template <typename E, typename S, typename = typename std::enable_if<std::is_arithmetic<S>::value>::type>
HIPACC_VEC_FN vec4<E> min(vec4<E> a, S s) {
  const E b = (E)s;
  return vec4<E>(vmin1(a.x, b), vmin1(a.y, b), vmin1(a.z, b), vmin1(a.w, b));
}
template <typename E, typename S, typename = typename std::enable_if<std::is_arithmetic<S>::value>::type>
HIPACC_VEC_FN vec4<E> min(S s, vec4<E> b) {
  const E a = (E)s;
  return vec4<E>(vmin1(a, b.x), vmin1(a, b.y), vmin1(a, b.z), vmin1(a, b.w));
}
template <typename E, typename S, typename = typename std::enable_if<std::is_arithmetic<S>::value>::type>
HIPACC_VEC_FN vec4<E> max(vec4<E> a, S s) {
  const E b = (E)s;
  return vec4<E>(vmax1(a.x, b), vmax1(a.y, b), vmax1(a.z, b), vmax1(a.w, b));
}
template <typename E, typename S, typename = typename std::enable_if<std::is_arithmetic<S>::value>::type>
HIPACC_VEC_FN vec4<E> max(S s, vec4<E> b) {
  const E a = (E)s;
  return vec4<E>(vmax1(a, b.x), vmax1(a, b.y), vmax1(a, b.z), vmax1(a, b.w));
}

template <typename E> HIPACC_VEC_FN vec4<E> abs(vec4<E> a) {
  auto ab = [](E v) -> E { return v < E(0) ? (E)(-v) : v; };
  return vec4<E>(ab(a.x), ab(a.y), ab(a.z), ab(a.w));
}

// Single-precision math of a float4, componentwise.
#define HIPACC_VEC_MATH1(F)                                                               \
  HIPACC_VEC_FN float4 F(float4 a) { return make_float4(::F(a.x), ::F(a.y), ::F(a.z), ::F(a.w)); }
HIPACC_VEC_MATH1(acosf)
HIPACC_VEC_MATH1(acoshf)
HIPACC_VEC_MATH1(asinf)
HIPACC_VEC_MATH1(asinhf)
HIPACC_VEC_MATH1(atanf)
HIPACC_VEC_MATH1(atanhf)
HIPACC_VEC_MATH1(cbrtf)
HIPACC_VEC_MATH1(ceilf)
HIPACC_VEC_MATH1(cosf)
HIPACC_VEC_MATH1(coshf)
HIPACC_VEC_MATH1(erff)
HIPACC_VEC_MATH1(erfcf)
HIPACC_VEC_MATH1(expf)
HIPACC_VEC_MATH1(exp2f)
HIPACC_VEC_MATH1(expm1f)
HIPACC_VEC_MATH1(fabsf)
HIPACC_VEC_MATH1(floorf)
HIPACC_VEC_MATH1(lgammaf)
HIPACC_VEC_MATH1(logf)
HIPACC_VEC_MATH1(log10f)
HIPACC_VEC_MATH1(log1pf)
HIPACC_VEC_MATH1(log2f)
HIPACC_VEC_MATH1(logbf)
HIPACC_VEC_MATH1(rintf)
HIPACC_VEC_MATH1(roundf)
HIPACC_VEC_MATH1(sinf)
HIPACC_VEC_MATH1(sinhf)
HIPACC_VEC_MATH1(sqrtf)
HIPACC_VEC_MATH1(tanf)
HIPACC_VEC_MATH1(tanhf)
HIPACC_VEC_MATH1(tgammaf)
HIPACC_VEC_MATH1(truncf)
#undef HIPACC_VEC_MATH1

}  // namespace math
}  // namespace hipacc

#endif  // HIPACC_HIP_VEC_HPP
