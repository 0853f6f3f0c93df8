// ring_u8.hip -- instantiations and launch geometry of the block row ring for
// 8-bit input pixels (hh_ring8.hpp): u8 -> u8 with a float sum, u8 -> i16 /
// u32 / i32 with an integer sum, square masks 3x3 and 5x5.
#include <cmath>

#include "hh_ring8.hpp"

namespace hh {

constexpr int kRing8Chunk = 8;   // rows per LDS chunk
constexpr int kRing8Sets = 1;    // register sets of staged rows in flight (edge strips)
constexpr int kRing8Depth = 2;   // chunks of direct loads in flight (interior strips)

// Time of an edge-strip row (16 per-pixel loads besides the 16-byte load)
// relative to an interior row.
constexpr float kRing8EdgeCost = 1.5f;

template <typename TO, bool ISUM, int K, bool FMA>
static int launch_ring8(const LocalArgs& a0, int seg_rows, hipStream_t s) {
  static const int slots = [] {
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu,
            reinterpret_cast<const void*>(
                &ring8_kernel<TO, ISUM, K, FMA, kRing8Chunk, kRing8Sets, kRing8Depth>),
            256, 0) != hipSuccess)
      return 0;
    return cus * (per_cu < 3 ? per_cu : 3);
  }();
  LocalArgs a = a0;
  stream_align(a.g, kRing8V, a.in_vec_ok, a.out_vec_ok);
  const long long strips = ceil_div(a.g.is_w, (64 - 2) * kRing8V);
  if (strips * a.g.is_h * a.g.batch > (1ll << 31))
    return set_error(HIPACC_ERR_UNSUPPORTED, "ring8 kernel: %lld strip rows",
                     strips * a.g.is_h * a.g.batch);
  const RingGeom rg = ring_geom(a.g, a.in_vec_ok, a.out_vec_ok, kRing8EdgeCost, 1, kRing8V, true,
                                kRing8Chunk, seg_rows, slots);
  hipLaunchKernelGGL((ring8_kernel<TO, ISUM, K, FMA, kRing8Chunk, kRing8Sets, kRing8Depth>),
                     dim3(rg.items), dim3(256), 0, s, a, rg);
  return hip_status(hipGetLastError(), "ring8_kernel launch");
}

// The ring beats the streaming kernel once most strips are interior: at
// 4096 columns (5 strips) 17 vs 20 us for a 3x3 Sobel u8 -> i16; at 1024
// (2 edge strips) the streaming kernel's short segments win, 6.7 vs 9.6 us.
bool ring8_pays(const LocalArgs& a) { return ceil_div(a.g.is_w, (64 - 2) * kRing8V) >= 4; }

int dispatch_ring8_local(const LocalArgs& a, int acc_is_float, int seg_rows, hipStream_t s) {
  const Geom& g = a.g;
  if (a.reduce != HIPACC_REDUCE_SUM || a.mask_w != a.mask_h || g.in_dtype != HIPACC_U8 ||
      a.wdev != nullptr || g.post_div != 0.0f)
    return HIPACC_ERR_UNSUPPORTED;
  const int bo = dtype_bytes(g.out_dtype);
  // byte offsets of the buffer accesses are 32-bit
  if ((long long)g.in_imgw >= (1ll << 31) || (long long)g.is_w * bo >= (1ll << 31))
    return HIPACC_ERR_UNSUPPORTED;
  const int k = a.mask_w;
  if (acc_is_float) {
    if (g.out_dtype != HIPACC_U8) return HIPACC_ERR_UNSUPPORTED;
#define HH_R8F(K)                                                                      \
  if (k == K) return a.fma ? launch_ring8<uint8_t, false, K, true>(a, seg_rows, s)     \
                           : launch_ring8<uint8_t, false, K, false>(a, seg_rows, s);
    HH_R8F(3) HH_R8F(5)
#undef HH_R8F
    return HIPACC_ERR_UNSUPPORTED;
  }
  // integer sums in float: exact while every partial sum stays below 2^24
  double wsum = 0.0;
  for (int t = 0; t < k * k; ++t) wsum += std::fabs((double)a.w.i[t]);
  if (255.0 * wsum >= 16777216.0) return HIPACC_ERR_UNSUPPORTED;
#define HH_R8I(TO, DT, K)                                                              \
  if (g.out_dtype == DT && k == K) return launch_ring8<TO, true, K, true>(a, seg_rows, s);
  HH_R8I(int16_t, HIPACC_I16, 3) HH_R8I(int16_t, HIPACC_I16, 5)
  HH_R8I(uint32_t, HIPACC_U32, 3) HH_R8I(uint32_t, HIPACC_U32, 5)
  HH_R8I(int32_t, HIPACC_I32, 3) HH_R8I(int32_t, HIPACC_I32, 5)
#undef HH_R8I
  return HIPACC_ERR_UNSUPPORTED;
}

}  // namespace hh
