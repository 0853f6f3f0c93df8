// hh_common.hpp -- pixel traits, boundary index maps and output conversions
// shared by every gfx950 kernel of libhipacc_gfx950.so.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/hipacc_hip.h"

#define HH_HD __host__ __device__ __forceinline__
// Pointer qualifier for read-only device data read with uniform indices: the
// AMDGPU constant address space, loaded with scalar (SMEM) loads.
#define HH_CONST __attribute__((address_space(4)))

namespace hh {

// ---------------------------------------------------------------------------
// Pixel type traits
// ---------------------------------------------------------------------------
template <typename T> struct dtype_of;
template <> struct dtype_of<uint8_t>  { static constexpr int value = HIPACC_U8;  };
template <> struct dtype_of<int8_t>   { static constexpr int value = HIPACC_I8;  };
template <> struct dtype_of<uint16_t> { static constexpr int value = HIPACC_U16; };
template <> struct dtype_of<int16_t>  { static constexpr int value = HIPACC_I16; };
template <> struct dtype_of<uint32_t> { static constexpr int value = HIPACC_U32; };
template <> struct dtype_of<int32_t>  { static constexpr int value = HIPACC_I32; };
template <> struct dtype_of<float>    { static constexpr int value = HIPACC_F32; };

// 4-wide vector pixels: four interleaved channels of the base scalar type.
HH_HD bool dtype_is_vec(int dt) { return dt >= HIPACC_U8X4 && dt <= HIPACC_F32X4; }
HH_HD int dtype_base(int dt) {
  switch (dt) {
    case HIPACC_U8X4: return HIPACC_U8;
    case HIPACC_I16X4: return HIPACC_I16;
    case HIPACC_I32X4: return HIPACC_I32;
    case HIPACC_F32X4: return HIPACC_F32;
    default: return dt;
  }
}
HH_HD int dtype_bytes(int dt) {
  const int lanes = dtype_is_vec(dt) ? 4 : 1;
  switch (dtype_base(dt)) {
    case HIPACC_U8: case HIPACC_I8: return 1 * lanes;
    case HIPACC_U16: case HIPACC_I16: return 2 * lanes;
    case HIPACC_U32: case HIPACC_I32: case HIPACC_F32: return 4 * lanes;
    default: return 0;
  }
}
HH_HD bool dtype_is_int(int dt) { return dt != HIPACC_F32; }

// Runtime-typed pixel load, converted to float or int32 (generic kernels).
HH_HD float load_f(const void* p, int dt, long long idx) {
  switch (dt) {
    case HIPACC_U8:  return (float)((const uint8_t*)p)[idx];
    case HIPACC_I8:  return (float)((const int8_t*)p)[idx];
    case HIPACC_U16: return (float)((const uint16_t*)p)[idx];
    case HIPACC_I16: return (float)((const int16_t*)p)[idx];
    case HIPACC_U32: return (float)((const uint32_t*)p)[idx];
    case HIPACC_I32: return (float)((const int32_t*)p)[idx];
    default:         return ((const float*)p)[idx];
  }
}
HH_HD int32_t load_i(const void* p, int dt, long long idx) {
  switch (dt) {
    case HIPACC_U8:  return (int32_t)((const uint8_t*)p)[idx];
    case HIPACC_I8:  return (int32_t)((const int8_t*)p)[idx];
    case HIPACC_U16: return (int32_t)((const uint16_t*)p)[idx];
    case HIPACC_I16: return (int32_t)((const int16_t*)p)[idx];
    case HIPACC_U32: return (int32_t)((const uint32_t*)p)[idx];
    case HIPACC_I32: return ((const int32_t*)p)[idx];
    default:         return (int32_t)((const float*)p)[idx];
  }
}

// ---------------------------------------------------------------------------
// Boundary handling: index maps of lib/AST/BorderHandling.cpp:41-120 over the
// accessor window [lo, hi).  REPEAT is the generated `while` form (periodic
// with period hi-lo).  MIRROR reflects once with the edge pixel repeated.
// Indices a single reflection cannot bring into the window (halo wider than
// the window) are left undefined by the reference; here they are clamped so
// that no access leaves the window.
// ---------------------------------------------------------------------------
HH_HD int bh_clamp(int i, int lo, int hi) { return i < lo ? lo : (i >= hi ? hi - 1 : i); }
HH_HD int bh_repeat(int i, int lo, int hi) {
  int n = hi - lo;
  int r = (i - lo) % n;
  if (r < 0) r += n;
  return lo + r;
}
HH_HD int bh_mirror(int i, int lo, int hi) {
  if (i < lo) i = lo + (lo - i - 1);
  if (i >= hi) i = hi - (i + 1 - hi);
  return bh_clamp(i, lo, hi);
}
// Maps i for a non-CONSTANT mode.  UNDEFINED reads the raw index; it is
// clamped to the allocation [0, ext) so that an out-of-image tap cannot fault.
HH_HD int bh_map(int mode, int i, int lo, int hi, int ext) {
  switch (mode) {
    case HIPACC_BOUNDARY_CLAMP:  return bh_clamp(i, lo, hi);
    case HIPACC_BOUNDARY_REPEAT: return bh_repeat(i, lo, hi);
    case HIPACC_BOUNDARY_MIRROR: return bh_mirror(i, lo, hi);
    default:                     return bh_clamp(i, 0, ext);
  }
}

// ---------------------------------------------------------------------------
// Output conversions (hipacc_cvt).  Float -> integer truncates toward zero
// with saturation (NaN -> 0): to the uint32 range for uint outputs, to the
// int32 range for every other integer output, which then wraps to its width.
// This is the conversion instruction of the GPU (v_cvt_u32_f32 /
// v_cvt_i32_f32); for every value the reference defines (in range of the
// output type) it is the C cast.
// ---------------------------------------------------------------------------
HH_HD int32_t f2i_sat(float v) {
#if defined(__HIP_DEVICE_COMPILE__)
  int32_t r;
  asm("v_cvt_i32_f32 %0, %1" : "=v"(r) : "v"(v));
  return r;
#else
  if (!(v == v)) return 0;
  if (v >= 2147483648.0f) return 2147483647;
  if (v <= -2147483648.0f) return (int32_t)0x80000000u;
  return (int32_t)v;
#endif
}
HH_HD uint32_t f2u_sat(float v) {
#if defined(__HIP_DEVICE_COMPILE__)
  uint32_t r;
  asm("v_cvt_u32_f32 %0, %1" : "=v"(r) : "v"(v));
  return r;
#else
  if (!(v == v) || v <= 0.0f) return 0u;
  if (v >= 4294967296.0f) return 0xffffffffu;
  return (uint32_t)v;
#endif
}
template <typename TO> struct cvt_out {
  HH_HD static TO from_f(float v) { return (TO)f2i_sat(v); }
  HH_HD static TO from_i(int32_t v) { return (TO)v; }
};
template <> struct cvt_out<uint32_t> {
  HH_HD static uint32_t from_f(float v) { return f2u_sat(v); }
  HH_HD static uint32_t from_i(int32_t v) { return (uint32_t)v; }
};
template <> struct cvt_out<float> {
  HH_HD static float from_f(float v) { return v; }
  HH_HD static float from_i(int32_t v) { return (float)v; }
};

template <typename TO>
HH_HD TO convert_f(float acc, int cvt, float post_div) {
  if (post_div != 0.0f) acc = acc / post_div;
  if (cvt == HIPACC_CVT_ROUND) acc = acc + 0.5f;
  return cvt_out<TO>::from_f(acc);
}
template <typename TO>
HH_HD TO convert_i(int32_t acc, int cvt, float post_div) {
  if (post_div != 0.0f) return convert_f<TO>((float)acc / post_div, cvt, 0.0f);
  return cvt_out<TO>::from_i(acc);
}

// Runtime-typed store (generic kernels).
HH_HD void store_f(void* p, int dt, long long idx, float acc, int cvt, float post_div) {
  switch (dt) {
    case HIPACC_U8:  ((uint8_t*)p)[idx]  = convert_f<uint8_t>(acc, cvt, post_div); break;
    case HIPACC_I8:  ((int8_t*)p)[idx]   = convert_f<int8_t>(acc, cvt, post_div); break;
    case HIPACC_U16: ((uint16_t*)p)[idx] = convert_f<uint16_t>(acc, cvt, post_div); break;
    case HIPACC_I16: ((int16_t*)p)[idx]  = convert_f<int16_t>(acc, cvt, post_div); break;
    case HIPACC_U32: ((uint32_t*)p)[idx] = convert_f<uint32_t>(acc, cvt, post_div); break;
    case HIPACC_I32: ((int32_t*)p)[idx]  = convert_f<int32_t>(acc, cvt, post_div); break;
    default:         ((float*)p)[idx]    = convert_f<float>(acc, cvt, post_div); break;
  }
}
HH_HD void store_i(void* p, int dt, long long idx, int32_t acc, int cvt, float post_div) {
  switch (dt) {
    case HIPACC_U8:  ((uint8_t*)p)[idx]  = convert_i<uint8_t>(acc, cvt, post_div); break;
    case HIPACC_I8:  ((int8_t*)p)[idx]   = convert_i<int8_t>(acc, cvt, post_div); break;
    case HIPACC_U16: ((uint16_t*)p)[idx] = convert_i<uint16_t>(acc, cvt, post_div); break;
    case HIPACC_I16: ((int16_t*)p)[idx]  = convert_i<int16_t>(acc, cvt, post_div); break;
    case HIPACC_U32: ((uint32_t*)p)[idx] = convert_i<uint32_t>(acc, cvt, post_div); break;
    case HIPACC_I32: ((int32_t*)p)[idx]  = convert_i<int32_t>(acc, cvt, post_div); break;
    default:         ((float*)p)[idx]    = convert_i<float>(acc, cvt, post_div); break;
  }
}

// Multiply-accumulate in the two float evaluation orders (see fp_contract).
template <bool FMA> HH_HD float madd(float w, float x, float acc) {
  if (FMA) return __builtin_fmaf(w, x, acc);
  float p = w * x;
  return acc + p;
}
template <bool FMA> HH_HD int32_t madd(int32_t w, int32_t x, int32_t acc) {
  return (int32_t)((uint32_t)acc + (uint32_t)w * (uint32_t)x);
}

// First tap of a sum: the product alone (the DSL seeds the result with the
// first tap, dsl/kernel.hpp:250); integer products wrap.
HH_HD float mul_first(float w, float x) { return w * x; }
HH_HD int32_t mul_first(int32_t w, int32_t x) { return (int32_t)((uint32_t)w * (uint32_t)x); }

// Reduction combine of hipacc::Reduce over the running value (dsl/kernel.hpp:
// 253-260): MIN/MAX take (tap, acc) in that argument order.
template <typename A> HH_HD A reduce_combine(int mode, A acc, A tap) {
  switch (mode) {
    case HIPACC_REDUCE_MIN: return tap < acc ? tap : acc;
    case HIPACC_REDUCE_MAX: return tap > acc ? tap : acc;
    case HIPACC_REDUCE_PROD: return acc * tap;
    default: return acc + tap;
  }
}
template <> HH_HD int32_t reduce_combine<int32_t>(int mode, int32_t acc, int32_t tap) {
  switch (mode) {
    case HIPACC_REDUCE_MIN: return tap < acc ? tap : acc;
    case HIPACC_REDUCE_MAX: return tap > acc ? tap : acc;
    case HIPACC_REDUCE_PROD: return (int32_t)((uint32_t)acc * (uint32_t)tap);
    default: return (int32_t)((uint32_t)acc + (uint32_t)tap);
  }
}

// Number of thread blocks helpers.
HH_HD int ceil_div(int a, int b) { return (a + b - 1) / b; }

}  // namespace hh
