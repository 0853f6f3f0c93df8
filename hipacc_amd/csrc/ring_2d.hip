// ring_2d.hip -- instantiations and launch geometry of the block row ring
// (hh_ring.hpp): f32 -> f32 weighted SUM, square masks 3x3 .. 9x9, and
// separable f32 masks 3 .. 13 taps.
#include <cmath>

#include "hh_ring.hpp"

namespace hh {

constexpr int kRingChunk = 8;   // rows per LDS chunk
constexpr int kRingSets = 2;    // register sets of staged rows in flight (edge strips)
// Chunks of direct loads in flight (interior strips): 2 for 7x7 and 9x9
// (4096^2: 40 -> 37.5 us and 64 -> 61 us), 1 for 3x3 and 5x5 (5x5 at depth
// 2: 27.0 vs 26.5 us).
constexpr int ring_depth(int k) { return k >= 7 ? 2 : 1; }

// Blocks resident on the whole device at once, at most 3 per CU (4 for
// 7x7): on a CU the oldest block is served first, so each further block
// lengthens the tail in which the youngest finishes alone (3 measured faster
// than 4 for 3x3 and 5x5, 4 faster for 7x7; 9x9 fits 3).
template <int K, bool FMA, bool WDEV>
static int ring_slots() {
  static const int slots = [] {
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu,
            reinterpret_cast<const void*>(
                &ring_kernel<K, FMA, WDEV, kRingChunk, kRingSets, ring_depth(K)>),
            256, 0) != hipSuccess)
      return 0;
    const int cap = K == 7 ? 4 : 3;
    return cus * (per_cu < cap ? per_cu : cap);
  }();
  return slots;
}

// Time of a row of an edge strip (per-pixel loads and stores) relative to a
// row of an interior strip, measured at 4096^2: the VALU work of the larger
// masks hides most of the extra memory instructions.
static float ring_edge_cost(int k) { return k <= 5 ? 1.4f : 1.15f; }

// Classifies the strips the way ring_kernel / ring_sep_kernel / ring8_kernel
// do (edge: the strip's staged columns leave the raw-readable region, or its
// stored columns are not all full vectors; v pixels per lane, hl halo lanes
// per side) and sizes the segments.  balance: the strips store the fewest
// whole lanes that still cover the width in the same number of strips (the
// lanes past them idle), instead of a narrow last strip.  seg_rows > 0 forces
// the interior segment height; 0 picks it so that one round of items fills
// `slots`.
RingGeom ring_geom(const Geom& g, int in_vec_ok, int out_vec_ok, float edge_cost, int hl, int v,
                   bool balance, int chunk, int seg_rows, int slots) {
  RingGeom rg;
  int sl = 64 - 2 * hl;                 // stored lanes per strip
  rg.strips = ceil_div(g.is_w, sl * v);
  if (balance) sl = ceil_div(ceil_div(g.is_w, rg.strips), v);
  rg.sw = sl * v;
  int lx = g.lo_x, hx = g.hi_x;
  if (g.boundary == HIPACC_BOUNDARY_UNDEFINED) {
    lx = 0;
    hx = g.in_imgw;
  }
  auto edge = [&](int strip) {
    const int xs = strip * rg.sw - hl * v, ix0 = g.in_ox + xs;
    return !(in_vec_ok && out_vec_ok && ix0 >= lx && ix0 + (sl + 2 * hl) * v <= hx &&
             xs + (sl + hl) * v <= g.is_w);
  };
  rg.nl = 0;
  while (rg.nl < rg.strips && edge(rg.nl)) ++rg.nl;
  rg.nr = 0;
  while (rg.nl + rg.nr < rg.strips && edge(rg.strips - 1 - rg.nr)) ++rg.nr;
  const int ne = rg.nl + rg.nr, nm = rg.strips - ne;
  const float ce = (nm > 0 && ne > 0) ? edge_cost : 1.0f;
  // chosen segment heights are whole chunks: a partial last chunk costs a
  // full step
  auto size = [&](int rows, bool whole) {
    int re = (int)((float)rows / ce);
    if (whole) {
      rows = ceil_div(rows, chunk) * chunk;
      re = ceil_div((int)((float)rows / ce), chunk) * chunk;
    }
    if (re < 8) re = 8;
    rg.seg_rows = rows < g.is_h ? rows : g.is_h;
    rg.seg_rows_e = re < g.is_h ? re : g.is_h;
    rg.segs = ceil_div(g.is_h, rg.seg_rows);
    rg.segs_e = ceil_div(g.is_h, rg.seg_rows_e);
    rg.per_img = nm * rg.segs + ne * rg.segs_e;
    rg.items = rg.per_img * g.batch;
  };
  if (seg_rows > 0) {
    size(seg_rows, false);
  } else {
    // interior-row units of work, spread over the slots; then taller
    // segments until the rounding leaves no item for a second round
    const double units = ((double)nm + ne * ce) * g.is_h * g.batch;
    int rows = slots > 0 ? (int)std::ceil(units / slots) : g.is_h;
    if (rows < 16) rows = 16;
    size(rows, true);
    for (int it = 0; it < 64 && slots > 0 && rg.items > slots && rows < g.is_h; ++it) {
      rows += rows / 32 + 1;
      size(rows, true);
    }
  }
  return rg;
}

template <int K, bool FMA, bool WDEV>
static int launch_ring(const LocalArgs& a0, int seg_rows, hipStream_t s) {
  LocalArgs a = a0;
  stream_align(a.g, kRingV, a.in_vec_ok, a.out_vec_ok);
  const long long strips = ceil_div(a.g.is_w, kRingStripW);
  if (strips * a.g.is_h * a.g.batch > (1ll << 31))
    return set_error(HIPACC_ERR_UNSUPPORTED, "ring kernel: %lld strip rows",
                     strips * a.g.is_h * a.g.batch);
  const RingGeom rg = ring_geom(a.g, a.in_vec_ok, a.out_vec_ok, ring_edge_cost(K), 1, kRingV,
                                false, kRingChunk, seg_rows, ring_slots<K, FMA, WDEV>());
  hipLaunchKernelGGL((ring_kernel<K, FMA, WDEV, kRingChunk, kRingSets, ring_depth(K)>),
                     dim3(rg.items), dim3(256), 0, s, a, rg);
  return hip_status(hipGetLastError(), "ring_kernel launch");
}

// Separable masks: 8-row chunks up to 9 taps, 16-row chunks for the larger
// vertical radii (up to 8), at most 3 resident blocks per CU, one register
// set of staged rows in flight.  At 4096^2: 13 taps f32 34.3 us (37.5 us
// with two sets); 5 taps f32 25.5 us and 7x7 box u8 -> u32 20 us (16-row
// chunks: 27.7 and 23.1 us).
template <int K> constexpr int sep_chunk() { return K <= 9 ? 8 : 16; }
constexpr int kRingSepSets = 1;

template <typename TI, typename TO, bool ISUM, int K, bool FMA>
static int launch_ring_sep(const SepArgs& a0, int seg_rows, hipStream_t s) {
  constexpr int HL = (K / 2 + kRingV - 1) / kRingV;
  static const int slots = [] {
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu,
            reinterpret_cast<const void*>(
                &ring_sep_kernel<TI, TO, K, K, FMA, ISUM, sep_chunk<K>(), kRingSepSets>),
            256, 0) != hipSuccess)
      return 0;
    return cus * (per_cu < 3 ? per_cu : 3);
  }();
  SepArgs a = a0;
  stream_align(a.g, kRingV, a.in_vec_ok, a.out_vec_ok);
  const long long strips = ceil_div(a.g.is_w, (64 - 2 * HL) * kRingV);
  if (strips * a.g.is_h * a.g.batch > (1ll << 31))
    return set_error(HIPACC_ERR_UNSUPPORTED, "ring_sep kernel: %lld strip rows",
                     strips * a.g.is_h * a.g.batch);
  const RingGeom rg = ring_geom(a.g, a.in_vec_ok, a.out_vec_ok, 1.3f, HL, kRingV, false,
                                sep_chunk<K>(), seg_rows, slots);
  hipLaunchKernelGGL((ring_sep_kernel<TI, TO, K, K, FMA, ISUM, sep_chunk<K>(), kRingSepSets>),
                     dim3(rg.items), dim3(256), 0, s, a, rg);
  return hip_status(hipGetLastError(), "ring_sep_kernel launch");
}

// f32 -> f32 (no rounding: the float output is the sum), u8 -> u8 (rounding
// as requested, saturating), and integer sums of u8 pixels into i16 / u32 /
// i32 (box sums and other integer masks), formed in float where every
// partial sum stays an exact integer: at most 255 * sum|row| per
// intermediate pixel (or the CONSTANT value), times sum|col|, below 2^24.
int dispatch_ring_sep(const SepArgs& a, int acc_is_float, int fma, int seg_rows, hipStream_t s) {
  const Geom& g = a.g;
  if (a.kx != a.ky || g.post_div != 0.0f) return HIPACC_ERR_UNSUPPORTED;
  const int bo = dtype_bytes(g.out_dtype);
  if ((long long)g.in_imgw * 4 >= (1ll << 31) || (long long)g.is_w * bo >= (1ll << 31))
    return HIPACC_ERR_UNSUPPORTED;
  if (!acc_is_float) {
    if (g.in_dtype != HIPACC_U8) return HIPACC_ERR_UNSUPPORTED;
    double sr = 0.0, sc = 0.0;
    for (int i = 0; i < a.kx; ++i) sr += std::fabs((double)a.rowi[i]);
    for (int i = 0; i < a.ky; ++i) sc += std::fabs((double)a.coli[i]);
    double hmax = 255.0 * sr;
    if (g.boundary == HIPACC_BOUNDARY_CONSTANT && std::fabs((double)g.const_i) > hmax)
      hmax = std::fabs((double)g.const_i);
    if (hmax * sc >= 16777216.0) return HIPACC_ERR_UNSUPPORTED;
#define HH_RSEPI(TO, DT, K)                                                            \
  if (g.out_dtype == DT && a.kx == K)                                                  \
    return launch_ring_sep<uint8_t, TO, true, K, true>(a, seg_rows, s);
    HH_RSEPI(uint32_t, HIPACC_U32, 3) HH_RSEPI(uint32_t, HIPACC_U32, 5)
    HH_RSEPI(uint32_t, HIPACC_U32, 7) HH_RSEPI(int32_t, HIPACC_I32, 3)
    HH_RSEPI(int32_t, HIPACC_I32, 5) HH_RSEPI(int32_t, HIPACC_I32, 7)
    HH_RSEPI(int16_t, HIPACC_I16, 3) HH_RSEPI(int16_t, HIPACC_I16, 5)
    HH_RSEPI(int16_t, HIPACC_I16, 7)
#undef HH_RSEPI
    return HIPACC_ERR_UNSUPPORTED;
  }
  const bool f32 = g.in_dtype == HIPACC_F32 && g.out_dtype == HIPACC_F32;
  const bool u8 = g.in_dtype == HIPACC_U8 && g.out_dtype == HIPACC_U8;
  if (!(f32 || u8) || (f32 && g.cvt == HIPACC_CVT_ROUND)) return HIPACC_ERR_UNSUPPORTED;
#define HH_RSEP(T, K)                                                                  \
  if (a.kx == K) return fma ? launch_ring_sep<T, T, false, K, true>(a, seg_rows, s)    \
                            : launch_ring_sep<T, T, false, K, false>(a, seg_rows, s);
  if (f32) { HH_RSEP(float, 3) HH_RSEP(float, 5) HH_RSEP(float, 7) HH_RSEP(float, 9) HH_RSEP(float, 13) }
  else { HH_RSEP(uint8_t, 3) HH_RSEP(uint8_t, 5) HH_RSEP(uint8_t, 7) HH_RSEP(uint8_t, 9) HH_RSEP(uint8_t, 13) }
#undef HH_RSEP
  return HIPACC_ERR_UNSUPPORTED;
}

int dispatch_ring_local(const LocalArgs& a, int seg_rows, hipStream_t s) {
  const Geom& g = a.g;
  if (a.reduce != HIPACC_REDUCE_SUM || a.mask_w != a.mask_h || g.in_dtype != HIPACC_F32 ||
      g.out_dtype != HIPACC_F32 || g.post_div != 0.0f || g.cvt == HIPACC_CVT_ROUND)
    return HIPACC_ERR_UNSUPPORTED;
  // byte offsets of the buffer accesses are 32-bit
  if ((long long)g.in_imgw * 4 >= (1ll << 31) || (long long)g.is_w * 4 >= (1ll << 31))
    return HIPACC_ERR_UNSUPPORTED;
  const bool wd = a.wdev != nullptr;
#define HH_RING(K)                                                                     \
  if (a.mask_w == K) {                                                                 \
    if (a.fma) return wd ? launch_ring<K, true, true>(a, seg_rows, s)                  \
                         : launch_ring<K, true, false>(a, seg_rows, s);                \
    return wd ? launch_ring<K, false, true>(a, seg_rows, s)                            \
              : launch_ring<K, false, false>(a, seg_rows, s);                          \
  }
  HH_RING(3) HH_RING(5) HH_RING(7) HH_RING(9)
#undef HH_RING
  return HIPACC_ERR_UNSUPPORTED;
}

}  // namespace hh
