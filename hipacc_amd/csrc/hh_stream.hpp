// hh_stream.hpp -- streaming local operators (weighted SUM 2D masks,
// separable masks, box sums) for the HBM-bound hot path.
//
// Each wave owns a vertical strip of the iteration space: lane l holds V
// adjacent columns (V = 4 or 8), so one row of the strip is one or two aligned
// vector loads per lane.  The wave walks its segment of rows top to bottom:
//  - rows are prefetched PF rows ahead into a register ring, so every lane
//    keeps PF rows of loads in flight without LDS or barriers;
//  - the horizontal halo comes from the neighbouring lanes through DPP
//    wave shifts (a VALU move, no LDS round trip); the HL lanes at either
//    edge only supply halo (they load but store nothing), so adjacent strips
//    overlap by 2*HL*V columns instead of issuing separate halo loads;
//  - every loaded row is added into the KY rolling accumulators of the output
//    rows whose window contains it, so each output pixel receives its taps in
//    row-major order (dy outer, dx inner), the order of lib/AST/
//    Convolution.cpp:397-435 and of the DSL's ElementIterator;
//  - waves whose rows and columns all lie inside the raw-readable region run
//    a lean loop: uniform row base + fixed lane offset, unconditional loads;
//    the waves at the window edges resolve the boundary mode through
//    per-lane column tables and a scalar row remap (lib/AST/
//    BorderHandling.cpp:41-120), still without divergent branches.
// Consecutive waves take vertically adjacent segments of one strip, and the
// block order is permuted so that blocks sharing an XCD (blockIdx % 8) take
// consecutive tiles: the halo rows two segments share hit the same L2.
#pragma once

#include "hh_tile.hpp"

namespace hh {

enum StreamKind { kStream2D = 0, kStreamSep = 1, kStreamBox = 2 };

struct StreamGeom {
  int strips, segs, seg_rows;  // tiling of one image
  int strip_w;                 // useful columns per strip
  int total_waves;
  int xcd_perm;                // permute blocks across the 8 XCDs
};

// Cross-lane moves by one lane through DPP wave shifts (gfx9 wave_shr:1 /
// wave_shl:1).  The lane without a source receives 0 (bound_ctrl), which
// spares the copy of the old value a DPP move with a kept value needs; the
// kernels never use what the edge lanes receive.
template <typename T> __device__ __forceinline__ T from_prev_lane(T v) {
  static_assert(sizeof(T) == 4, "32-bit lane values");
  const int x = __builtin_bit_cast(int, v);
  return __builtin_bit_cast(T, __builtin_amdgcn_mov_dpp(x, 0x138, 0xf, 0xf, true));
}
template <typename T> __device__ __forceinline__ T from_next_lane(T v) {
  static_assert(sizeof(T) == 4, "32-bit lane values");
  const int x = __builtin_bit_cast(int, v);
  return __builtin_bit_cast(T, __builtin_amdgcn_mov_dpp(x, 0x130, 0xf, 0xf, true));
}
template <typename T> __device__ __forceinline__ T lane_shift(T v, int d) {
  // value of lane (l + d); d is a compile-time constant after unrolling
  for (int k = 0; k < d; ++k) v = from_next_lane(v);
  for (int k = 0; k < -d; ++k) v = from_prev_lane(v);
  return v;
}

// Per-lane column sources for waves that touch a window edge: the index each
// of the lane's pixels reads (boundary-mapped, always a valid address) and
// whether it reads the CONSTANT value instead.
template <int V> struct LaneCols {
  int x[V];
  bool cst[V];
};

template <int V> __device__ __forceinline__ LaneCols<V> lane_cols(const Geom& g, int ix) {
  LaneCols<V> c;
  const bool cmode = g.boundary == HIPACC_BOUNDARY_CONSTANT;
#pragma unroll
  for (int e = 0; e < V; ++e) {
    const int x = ix + e;
    if (cmode) {
      c.cst[e] = x < g.lo_x || x >= g.hi_x;
      c.x[e] = c.cst[e] ? g.lo_x : x;
    } else {
      c.cst[e] = false;
      c.x[e] = bh_map(g.boundary, x, g.lo_x, g.hi_x, g.in_imgw);
    }
  }
  return c;
}

// Row source of input row iy (absolute) for edge waves: the mapped row and
// whether the whole row reads the CONSTANT value.
__device__ __forceinline__ int row_src(const Geom& g, int iy, bool& cst) {
  if (g.boundary == HIPACC_BOUNDARY_CONSTANT) {
    cst = iy < g.lo_y || iy >= g.hi_y;
    return cst ? g.lo_y : iy;
  }
  cst = false;
  return bh_map(g.boundary, iy, g.lo_y, g.hi_y, g.in_imgh);
}

template <typename ArgsT> struct StreamTaps;
template <> struct StreamTaps<LocalArgs> {
  template <typename ACC, bool WDEV>
  __device__ __forceinline__ static ACC w2d(const LocalArgs& a, int t) {
    // Device weights are read through the constant address space: uniform
    // scalar loads into SGPRs (a generic pointer would make them per-lane
    // vector loads held in VGPRs, since the output stores might alias).
    if constexpr (WDEV) return (ACC)((const HH_CONST float*)a.wdev)[t];
    else if constexpr (std::is_same<ACC, float>::value) return a.w.f[t];
    else return (ACC)a.w.i[t];
  }
};

// V pixels of T: loaded in 16-byte pieces (or one smaller word).
template <typename T, int V>
__device__ __forceinline__ void load_px(T (&v)[V], const T* p) {
  constexpr int B = V * (int)sizeof(T);
  if constexpr (B >= 16) {
#pragma unroll
    for (int k = 0; k < B / 16; ++k) {
      const uint4 w = reinterpret_cast<const uint4*>(p)[k];
      __builtin_memcpy(reinterpret_cast<char*>(v) + 16 * k, &w, 16);
    }
  } else {
    using W = typename word_of<B>::type;
    const W w = *reinterpret_cast<const W*>(p);
    __builtin_memcpy(v, &w, B);
  }
}

// Everything one wave needs while streaming its segment.
template <typename TI, typename TO, int V>
struct StreamCtx {
  const TI* in;        // image plane
  TO* out;
  int ix;              // absolute input column of the lane's pixel 0
  int xl;              // iteration-space column of the lane's pixel 0
  int iy0;             // absolute input row of streamed row 0
  int y0;              // iteration-space row of output row 0
  int nj;              // rows streamed
  bool store_lane;
  LaneCols<V> cols;
};

// Edge-wave row load: mapped row, per-lane mapped columns, constant select.
template <typename TI, int V>
__device__ __forceinline__ void edge_row_load(TI (&v)[V], const Geom& g, const TI* __restrict__ in,
                                              const LaneCols<V>& cols, int iy) {
  bool rc;
  const int y = row_src(g, iy, rc);
  const TI* row = in + (long long)y * g.in_stride;
  const TI cv = const_pixel<TI>(g);
#pragma unroll
  for (int e = 0; e < V; ++e) {
    const TI t = row[cols.x[e]];
    v[e] = (rc || cols.cst[e]) ? cv : t;
  }
}

// One streamed row: widen, gather the halo from the neighbour lanes, and add
// the row into the accumulators of the KY output rows it belongs to.  `slot0`
// is the accumulator of output row j (= j mod KY); `cst_row` (separable,
// CONSTANT) makes the intermediate row the constant.  Only output rows j-dyi
// with mindy <= dyi <= maxdy exist: maxdy < KY-1 in the first rows of a
// segment (a constant after unrolling, so the skipped taps cost nothing),
// mindy > 0 in its last rows, past the segment's last output.
template <typename TI, typename ACC, int V, int KX, int KY, bool FMA, int KIND, bool WDEV,
          typename ArgsT>
__device__ __forceinline__ void stream_accumulate(const ArgsT& a, const TI (&raw)[V],
                                                  ACC (&acc)[KY][V], int slot0, bool cst_row,
                                                  int maxdy, int mindy) {
  constexpr int RX = KX / 2;
  constexpr int NWIN = V + 2 * RX;
  constexpr bool IS_F = std::is_same<ACC, float>::value;
  ACC cur[V];
#pragma unroll
  for (int e = 0; e < V; ++e) cur[e] = (ACC)raw[e];
  ACC win[NWIN];
#pragma unroll
  for (int wi = 0; wi < NWIN; ++wi) {
    const int rel = wi - RX;
    const int d = rel >= 0 ? rel / V : -((-rel + V - 1) / V);
    const int e = rel - d * V;
    win[wi] = d == 0 ? cur[e] : lane_shift(cur[e], d);
  }
  if constexpr (KIND == kStream2D) {
#pragma unroll
    for (int dyi = 0; dyi < KY; ++dyi) {
      if (dyi > maxdy) break;
      if (dyi < mindy) continue;
      const int s = (slot0 - dyi + KY) % KY;
#pragma unroll
      for (int dx = 0; dx < KX; ++dx) {
        const ACC w = StreamTaps<ArgsT>::template w2d<ACC, WDEV>(a, dyi * KX + dx);
#pragma unroll
        for (int e = 0; e < V; ++e) {
          // the first tap of an output starts its sum: round(w*x) == fma(w, x, 0)
          if (dyi == 0 && dx == 0) acc[s][e] = mul_first(w, win[e]);
          else acc[s][e] = madd<FMA>(w, win[e + dx], acc[s][e]);
        }
      }
    }
  } else {
    ACC h[V];
    if constexpr (KIND == kStreamBox) {
      ACC s0 = ACC(0);
#pragma unroll
      for (int dx = 0; dx < KX; ++dx) s0 = (ACC)((uint32_t)s0 + (uint32_t)win[dx]);
      h[0] = s0;
#pragma unroll
      for (int e = 1; e < V; ++e)
        h[e] = (ACC)((uint32_t)h[e - 1] + (uint32_t)win[e + KX - 1] - (uint32_t)win[e - 1]);
    } else {
#pragma unroll
      for (int e = 0; e < V; ++e) {
        ACC s0 = ACC(0);
#pragma unroll
        for (int dx = 0; dx < KX; ++dx) {
          const ACC w = IS_F ? (ACC)a.rowf[dx] : (ACC)a.rowi[dx];
          s0 = dx == 0 ? mul_first(w, win[e]) : madd<FMA>(w, win[e + dx], s0);
        }
        h[e] = s0;
      }
    }
    if (cst_row) {
      const ACC cv = IS_F ? (ACC)a.g.const_f : (ACC)a.g.const_i;
#pragma unroll
      for (int e = 0; e < V; ++e) h[e] = cv;
    }
#pragma unroll
    for (int dyi = 0; dyi < KY; ++dyi) {
      if (dyi > maxdy) break;
      if (dyi < mindy) continue;
      const int s = (slot0 - dyi + KY) % KY;
      if constexpr (KIND == kStreamBox) {
#pragma unroll
        for (int e = 0; e < V; ++e)
          acc[s][e] = dyi == 0 ? h[e] : (ACC)((uint32_t)acc[s][e] + (uint32_t)h[e]);
      } else {
        const ACC w = IS_F ? (ACC)a.colf[dyi] : (ACC)a.coli[dyi];
#pragma unroll
        for (int e = 0; e < V; ++e)
          acc[s][e] = dyi == 0 ? mul_first(w, h[e]) : madd<FMA>(w, h[e], acc[s][e]);
      }
    }
  }
}

// Output row store of one lane (row in range; columns checked here).  The
// streaming variants are only chosen without post_div and, for float
// outputs, without rounding, so the conversion needs no division; full
// vectors are written with nontemporal stores (the output is not re-read).
template <typename TO, typename ACC, int V>
__device__ __forceinline__ void stream_store(TO* __restrict__ orow, const Geom& g, bool full,
                                             bool any, int xl, float rnd, const ACC (&acc)[V]) {
  if (!any) return;
  TO vals[V];
#pragma unroll
  for (int e = 0; e < V; ++e) {
    if constexpr (std::is_same<TO, float>::value)
      vals[e] = (float)acc[e];
    else if constexpr (std::is_same<ACC, float>::value)
      vals[e] = cvt_out<TO>::from_f((float)acc[e] + rnd);
    else
      vals[e] = cvt_out<TO>::from_i((int32_t)acc[e]);
  }
  TO* dst = orow + xl;
  if (full) {
    store_vec<TO, V>(dst, vals);
  } else {
#pragma unroll
    for (int e = 0; e < V; ++e)
      if (xl + e < g.is_w) dst[e] = vals[e];
  }
}

// One block of PF streamed rows starting at row jb (the ring index u and the
// accumulator slot are compile-time constants inside a block).  WARM: the
// first block of the segment, whose row u only reaches outputs 0..u.
template <typename TI, typename TO, typename ACC, int V, int KX, int KY, int PF, bool FMA,
          int KIND, bool WDEV, bool FAST, bool WARM, bool TAIL, typename ArgsT>
__device__ __forceinline__ void stream_block(const ArgsT& a, const StreamCtx<TI, TO, V>& c,
                                             TI (&ring)[PF][V], ACC (&acc)[KY][V], const TI* base,
                                             TO* orow, bool full, bool any, float rnd, int jb) {
  const Geom& g = a.g;
  const bool cmode_sep = KIND != kStream2D && !FAST && g.boundary == HIPACC_BOUNDARY_CONSTANT;
  // every refill of this block is in range (always for the non-tail blocks)
  const bool full_blk = !TAIL || jb + 2 * PF <= c.nj;
#pragma unroll
  for (int u = 0; u < PF; ++u) {
    const int j = jb + u;
    if (!full_blk && j >= c.nj) break;
    TI raw[V];
#pragma unroll
    for (int e = 0; e < V; ++e) raw[e] = ring[u][e];
    const int r = j + PF;
    if (full_blk || r < c.nj) {
      if constexpr (FAST) load_px<TI, V>(ring[u], base + (long long)r * g.in_stride);
      else edge_row_load<TI, V>(ring[u], g, c.in, c.cols, c.iy0 + r);
    }
    bool cst_row = false;
    if (cmode_sep) {
      const int iy = c.iy0 + j;
      cst_row = iy < g.lo_y || iy >= g.hi_y;
    }
    // The taps of outputs past the segment are skipped for the larger masks,
    // where they are a fifth of the work or more; for small masks the extra
    // branches cost more than they save.
    constexpr bool SKIP_TAIL = TAIL && KY >= 7;
    stream_accumulate<TI, ACC, V, KX, KY, FMA, KIND, WDEV>(a, raw, acc, u % KY, cst_row,
                                                           WARM ? u : KY - 1,
                                                           SKIP_TAIL ? j - c.nj + KY : 0);
    if (!WARM || u >= KY - 1)
      stream_store<TO, ACC, V>(orow + (long long)j * g.out_stride, g, full, any, c.xl, rnd,
                               acc[(u % KY + 1) % KY]);
  }
}

// Streams the wave's segment: prefetch PF rows, then blocks of PF rows.  The
// first KY-1 rows complete no output.  Blocks whose refills all stay inside
// the segment load unconditionally; the last blocks skip the refills past the
// segment's last row (a wave-uniform branch).
template <typename TI, typename TO, typename ACC, int V, int KX, int KY, int PF, bool FMA,
          int KIND, bool WDEV, bool FAST, typename ArgsT>
__device__ __forceinline__ void stream_rows(const ArgsT& a, const StreamCtx<TI, TO, V>& c) {
  const Geom& g = a.g;
  const bool full = c.store_lane && a.out_vec_ok && c.xl + V <= g.is_w;
  const bool any = c.store_lane && c.xl < g.is_w;
  const float rnd = g.cvt == HIPACC_CVT_ROUND ? 0.5f : 0.0f;
  TO* orow = c.out + (long long)(g.out_oy + c.y0 - (KY - 1)) * g.out_stride + g.out_ox;
  // FAST rows: wave-uniform row base plus this lane's fixed offset
  const TI* base = c.in + (long long)c.iy0 * g.in_stride + c.ix;
  TI ring[PF][V];
  ACC acc[KY][V];
#pragma unroll
  for (int p = 0; p < PF; ++p) {
    const int r = min(p, c.nj - 1);
    if constexpr (FAST) load_px<TI, V>(ring[p], base + (long long)r * g.in_stride);
    else edge_row_load<TI, V>(ring[p], g, c.in, c.cols, c.iy0 + r);
  }
  stream_block<TI, TO, ACC, V, KX, KY, PF, FMA, KIND, WDEV, FAST, true, true>(
      a, c, ring, acc, base, orow, full, any, rnd, 0);
  if constexpr (KY >= 7) {
    // full blocks, then the tail blocks that skip the taps past the segment
    int jb = PF;
    for (; jb + 2 * PF <= c.nj; jb += PF)
      stream_block<TI, TO, ACC, V, KX, KY, PF, FMA, KIND, WDEV, FAST, false, false>(
          a, c, ring, acc, base, orow, full, any, rnd, jb);
    for (; jb < c.nj; jb += PF)
      stream_block<TI, TO, ACC, V, KX, KY, PF, FMA, KIND, WDEV, FAST, false, true>(
          a, c, ring, acc, base, orow, full, any, rnd, jb);
  } else {
    for (int jb = PF; jb < c.nj; jb += PF)
      stream_block<TI, TO, ACC, V, KX, KY, PF, FMA, KIND, WDEV, FAST, false, true>(
          a, c, ring, acc, base, orow, full, any, rnd, jb);
  }
}

template <typename TI, typename TO, typename ACC, int V, int KX, int KY, int PF, bool FMA,
          int KIND, bool WDEV, typename ArgsT>
__global__ __launch_bounds__(256) void stream_kernel(const ArgsT a, const StreamGeom sg) {
  constexpr int RX = KX / 2, RY = KY / 2;
  constexpr int HL = (RX + V - 1) / V;          // halo lanes on each side
  static_assert(PF % KY == 0 && PF >= KY, "prefetch ring must be a multiple of KY");

  const Geom& g = a.g;
  const int lane = threadIdx.x & 63;
  int blk = blockIdx.x;
  if (sg.xcd_perm) {
    const int per = gridDim.x >> 3;
    blk = (blk & 7) * per + (blk >> 3);
  }
  const int wv = blk * 4 + (threadIdx.x >> 6);
  if (wv >= sg.total_waves) return;
  const int per_img = sg.strips * sg.segs;
  const int b = wv / per_img;
  const int t = wv - b * per_img;
  const int strip = t / sg.segs;
  const int seg = t - strip * sg.segs;

  StreamCtx<TI, TO, V> c;
  c.in = static_cast<const TI*>(g.in) + (long long)b * g.in_plane;
  c.out = static_cast<TO*>(g.out) + (long long)b * g.out_plane;
  c.xl = strip * sg.strip_w - HL * V + lane * V;
  c.ix = g.in_ox + c.xl;
  c.y0 = seg * sg.seg_rows;
  const int nout = min(sg.seg_rows, g.is_h - c.y0);
  c.nj = nout + KY - 1;
  c.iy0 = g.in_oy + c.y0 - RY;
  c.store_lane = lane >= HL && lane < 64 - HL;
  int lx, hx, ly, hy;
  raw_region(g, lx, hx, ly, hy);
  // The wave reads raw rows iy0 .. iy0+nj-1 and the columns of all 64 lanes.
  const int wx0 = g.in_ox + strip * sg.strip_w - HL * V;
  const bool fast = a.in_vec_ok && wx0 >= lx && wx0 + 64 * V <= hx && c.iy0 >= ly &&
                    c.iy0 + c.nj <= hy;
  if (fast) {
    stream_rows<TI, TO, ACC, V, KX, KY, PF, FMA, KIND, WDEV, true>(a, c);
  } else {
    c.cols = lane_cols<V>(g, c.ix);
    stream_rows<TI, TO, ACC, V, KX, KY, PF, FMA, KIND, WDEV, false>(a, c);
  }
}

// Alignment the streaming kernels need: V-pixel loads and stores.
static inline void stream_align(const Geom& g, int v, int& in_ok, int& out_ok) {
  const int ebi = dtype_bytes(g.in_dtype), ebo = dtype_bytes(g.out_dtype);
  const int ai = v * ebi < 16 ? v * ebi : 16, ao = v * ebo < 16 ? v * ebo : 16;
  const uintptr_t pi = reinterpret_cast<uintptr_t>(g.in), po = reinterpret_cast<uintptr_t>(g.out);
  in_ok = (pi % ai) == 0 && ((long long)g.in_stride * ebi) % ai == 0 &&
          ((long long)g.in_ox * ebi) % ai == 0 && (g.batch <= 1 || (g.in_plane * ebi) % ai == 0);
  out_ok = (po % ao) == 0 && ((long long)g.out_stride * ebo) % ao == 0 &&
           ((long long)g.out_ox * ebo) % ao == 0 &&
           (g.batch <= 1 || (g.out_plane * ebo) % ao == 0);
}

static inline StreamGeom stream_geom(const Geom& g, int rx, int v, int seg_rows) {
  StreamGeom sg;
  const int hl = (rx + v - 1) / v;
  sg.strip_w = (64 - 2 * hl) * v;
  sg.strips = ceil_div(g.is_w, sg.strip_w);
  if (seg_rows <= 0) {
    // 12-row segments, shorter on small images until there are about 2560
    // waves (10 per CU): a wave walks its rows one after another, so few
    // long segments leave the launch latency-bound (1024^2 u8: 2-row
    // segments 8.8 us, 12-row 13 us)
    const long long wave_rows = (long long)sg.strips * g.is_h * g.batch;
    seg_rows = (int)((wave_rows + 2559) / 2560);
    seg_rows = seg_rows < 2 ? 2 : seg_rows > 12 ? 12 : seg_rows;
  }
  sg.seg_rows = seg_rows;
  sg.segs = ceil_div(g.is_h, sg.seg_rows);
  sg.total_waves = sg.strips * sg.segs * g.batch;
  sg.xcd_perm = 0;
  return sg;
}

static inline dim3 stream_grid(StreamGeom& sg) {
  int blocks = ceil_div(sg.total_waves, 4);
  if (blocks >= 64) {
    blocks = (blocks + 7) / 8 * 8;
    sg.xcd_perm = 1;
  }
  return dim3(blocks);
}

template <typename TI, typename TO, typename ACC, int V, int KX, int KY, int PF, bool FMA,
          int KIND, bool WDEV, typename ArgsT>
static int launch_stream(const ArgsT& a0, int seg_rows, hipStream_t s) {
  if (a0.g.post_div != 0.0f ||
      (std::is_same<TO, float>::value && a0.g.cvt == HIPACC_CVT_ROUND))
    return HIPACC_ERR_UNSUPPORTED;
  ArgsT a = a0;
  stream_align(a.g, V, a.in_vec_ok, a.out_vec_ok);
  StreamGeom sg = stream_geom(a.g, KX / 2, V, seg_rows);
  const dim3 grid = stream_grid(sg);
  hipLaunchKernelGGL((stream_kernel<TI, TO, ACC, V, KX, KY, PF, FMA, KIND, WDEV, ArgsT>), grid,
                     dim3(256), 0, s, a, sg);
  return hip_status(hipGetLastError(), "stream_kernel launch");
}

// Prefetch depth in rows: a multiple of KY (at least 6).
#define HH_PF(K) ((K) >= 5 ? (K) : 6)

}  // namespace hh
