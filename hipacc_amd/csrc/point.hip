// point.hip -- point operators (Kernel::execute with a per-pixel output()).
//
// Each thread produces 4 horizontally adjacent pixels of one row; rows of the
// iteration space map to blockIdx.y so no per-pixel division is needed.  Loads
// and stores are issued as one vector when the windows are suitably aligned.
#include "hh_tile.hpp"
#include "hh_point.hpp"

namespace hh {

template <typename T> __device__ __forceinline__ void load4(const T* p, bool vec, T (&v)[4], int n) {
  if (vec && n == 4) {
    using W = typename word_of<4 * sizeof(T)>::type;
    W w = *reinterpret_cast<const W*>(p);
    __builtin_memcpy(v, &w, sizeof(W));
  } else {
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = k < n ? p[k] : T(0);
  }
}

// Integer inputs: int arithmetic except AFFINE (float) and MAGNITUDE (int
// division by the norm, int sum of squares, float sqrt and clamp).
template <typename T0, typename T1, typename TO, bool FLOAT_MATH>
__global__ __launch_bounds__(256) void point_kernel(const PointArgs a) {
  const int x = (blockIdx.x * 256 + threadIdx.x) * 4;
  const int y = blockIdx.y;
  if (x >= a.w) return;
  const int n = a.w - x < 4 ? a.w - x : 4;
  const long long b = blockIdx.z;
  const T0* p0 = static_cast<const T0*>(a.in0) + b * a.plane0 + (long long)(a.oy0 + y) * a.stride0 + a.ox0 + x;
  T0 v0[4];
  T1 v1[4] = {T1(0), T1(0), T1(0), T1(0)};
  load4<T0>(p0, a.vec_ok, v0, n);
  if (a.in1) {
    const T1* p1 = static_cast<const T1*>(a.in1) + b * a.plane1 + (long long)(a.oy1 + y) * a.stride1 + a.ox1 + x;
    load4<T1>(p1, a.vec_ok, v1, n);
  }
  TO r[4];
#pragma unroll
  for (int k = 0; k < 4; ++k)
    r[k] = point_pixel<TO, FLOAT_MATH>(a.op, (float)v0[k], (float)v1[k], (int32_t)v0[k],
                                       (int32_t)v1[k], a.alpha, a.beta, a.cvt);
  TO* po = static_cast<TO*>(a.out) + b * a.plane_out + (long long)(a.oyo + y) * a.stride_out + a.oxo + x;
  if (a.vec_ok && n == 4) {
    store_vec<TO, 4>(po, r);
  } else {
    for (int k = 0; k < n; ++k) po[k] = r[k];
  }
}

template <typename T0, typename T1, typename TO>
static int launch_point_t(const PointArgs& a, hipStream_t s) {
  dim3 grid(ceil_div(a.w, 1024), a.h, a.batch);
  constexpr bool FM = std::is_same<T0, float>::value || std::is_same<T1, float>::value;
  hipLaunchKernelGGL((point_kernel<T0, T1, TO, FM>), grid, dim3(256), 0, s, a);
  return hip_status(hipGetLastError(), "point_kernel launch");
}

template <typename T0, typename T1>
static int launch_point_out(const PointArgs& a, int dto, hipStream_t s) {
  switch (dto) {
    case HIPACC_U8: return launch_point_t<T0, T1, uint8_t>(a, s);
    case HIPACC_I8: return launch_point_t<T0, T1, int8_t>(a, s);
    case HIPACC_U16: return launch_point_t<T0, T1, uint16_t>(a, s);
    case HIPACC_I16: return launch_point_t<T0, T1, int16_t>(a, s);
    case HIPACC_U32: return launch_point_t<T0, T1, uint32_t>(a, s);
    case HIPACC_I32: return launch_point_t<T0, T1, int32_t>(a, s);
    case HIPACC_F32: return launch_point_t<T0, T1, float>(a, s);
  }
  return HIPACC_ERR_INVALID_ARG;
}

// Both inputs share one pixel type (the second input, when present, must
// have the first one's type).
int launch_point(const PointArgs& a, int dt0, int dt1, int dto, hipStream_t s) {
  if (a.in1 && dt1 != dt0) return set_error(HIPACC_ERR_UNSUPPORTED, "point: inputs of different pixel types");
  switch (dt0) {
    case HIPACC_U8: return launch_point_out<uint8_t, uint8_t>(a, dto, s);
    case HIPACC_I8: return launch_point_out<int8_t, int8_t>(a, dto, s);
    case HIPACC_U16: return launch_point_out<uint16_t, uint16_t>(a, dto, s);
    case HIPACC_I16: return launch_point_out<int16_t, int16_t>(a, dto, s);
    case HIPACC_U32: return launch_point_out<uint32_t, uint32_t>(a, dto, s);
    case HIPACC_I32: return launch_point_out<int32_t, int32_t>(a, dto, s);
    case HIPACC_F32: return launch_point_out<float, float>(a, dto, s);
  }
  return HIPACC_ERR_INVALID_ARG;
}

}  // namespace hh
