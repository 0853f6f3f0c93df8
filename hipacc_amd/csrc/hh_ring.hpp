// hh_ring.hpp -- block row ring for the 2D weighted SUM, f32 -> f32.
//
// A block of 4 waves owns a strip of 256 staged columns (lane l holds 4
// adjacent pixels; lanes 0 and 63 only supply the horizontal halo, so 248
// columns are stored) and a segment of output rows.  The input rows of the
// segment pass once through a ring of three CH-row chunks in LDS: the four
// waves stage the rows of a chunk together (CH/4 rows each), and each wave
// then computes CH/4 output rows of the chunk from the LDS rows.  The
// vertical halo of the mask is therefore read once per block segment (KY-1
// rows per segment of ~70 rows), not once per wave segment as in the
// register streaming kernel of hh_stream.hpp.  NS register sets keep the
// loads of the next NS chunks in flight while a chunk is computed.  The
// segments are sized so that every block of the launch is resident at once
// (ring_2d.hip).
//
// Each output pixel receives its taps in row-major order (dy outer, dx
// inner), the order of lib/AST/Convolution.cpp:397-435, with the first tap
// a plain product -- the results equal the streaming and tile kernels' bit
// for bit.  The boundary modes follow lib/AST/BorderHandling.cpp:41-120:
// rows are remapped per staged row (a scalar), columns per lane in the
// strips that touch the window's left or right edge, where only the lanes
// outside the raw-readable columns load pixel by pixel.
//
// The main loop holds no divergent branch and no conditional memory
// operation, so the compiler's memory wait counts stay exact: rows past the
// segment are clamped on load, CONSTANT pixels are selected after the load,
// and stores go through buffer resources whose range check drops the halo
// lanes, the columns past the window and the rows past the segment.
#pragma once

#include "hh_stream.hpp"

namespace hh {

constexpr int kRingV = 4;                      // pixels per lane
constexpr int kRingStripW = (64 - 2) * kRingV;  // stored columns per strip
constexpr int kRingDrop = (int)0x80000000;      // buffer offset past every range: access dropped

// Strips [0, nl) and [strips - nr, strips) take the per-pixel (edge) path;
// their segments are shorter, sized so that every item of the launch
// carries about the same time.
struct RingGeom {
  int strips, nl, nr;
  int sw;                    // stored columns per strip
  int segs, seg_rows;        // interior strips
  int segs_e, seg_rows_e;    // edge strips
  int per_img;               // items (strip segments) per image
  int items;                 // per_img * batch
};

typedef unsigned int ring_u4 __attribute__((ext_vector_type(4)));

// Strip classification and segment sizes of a launch (ring_2d.hip).
RingGeom ring_geom(const Geom& g, int in_vec_ok, int out_vec_ok, float edge_cost, int hl, int v,
                   bool balance, int chunk, int seg_rows, int slots);

// Buffer resource over `bytes` bytes from `base` (raw buffer, no stride):
// accesses at offsets past `bytes` are dropped (stores) or read 0 (loads).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t ring_rsrc(const void* base, int bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, bytes, 0x00020000);
}

// Staged rows in flight.  With split edge loads a row is loaded twice: one
// 16-byte load for the lanes whose pixels are all raw-readable, one 4-byte
// load per pixel for the others (the other load of each lane is dropped and
// reads 0); the two are merged when the row is written to LDS.
template <bool SPLIT, int QR> struct RingStage {
  float4 r[QR];
  int rc;          // bit i: staged row i is outside the window in CONSTANT mode
};
template <int QR> struct RingStage<true, QR> {
  float4 r[QR];
  float4 p[QR];
  int rc;
};

// Waits until at most N vector memory operations of this wave are
// outstanding (they complete in issue order), and a workgroup barrier; both
// are compiler barriers for memory operations.  The barrier carries no wait
// of its own, so direct loads stay in flight across it.
template <int N> __device__ __forceinline__ void ring_wait_vm() {
  static_assert(N >= 0 && N < 64, "vmcnt range");
  asm volatile("s_waitcnt vmcnt(%0)" ::"n"(N) : "memory");
}
__device__ __forceinline__ void ring_barrier() { asm volatile("s_barrier" ::: "memory"); }

__device__ __forceinline__ float4 ring_merge(const float4& a, const float4& b) {
  // one of the two is all zero bits: OR keeps the other bit for bit
  return make_float4(__builtin_bit_cast(float, __builtin_bit_cast(unsigned, a.x) | __builtin_bit_cast(unsigned, b.x)),
                     __builtin_bit_cast(float, __builtin_bit_cast(unsigned, a.y) | __builtin_bit_cast(unsigned, b.y)),
                     __builtin_bit_cast(float, __builtin_bit_cast(unsigned, a.z) | __builtin_bit_cast(unsigned, b.z)),
                     __builtin_bit_cast(float, __builtin_bit_cast(unsigned, a.w) | __builtin_bit_cast(unsigned, b.w)));
}

// EDGE: the strip's columns leave the raw-readable region, or its stored
// columns are not all full vectors (per-lane column remap, per-pixel loads
// and stores where needed); otherwise one 16-byte load and store per lane
// and row.
// The block computes output rows [y0, y0 + nout) of its strip; xl is the
// iteration-space column of the lane's first pixel.  CH rows per LDS chunk.
// Interior strips stage their rows with direct global->LDS loads, ND chunks
// ahead, through a ring of ND+2 chunks; edge strips stage through NS
// register sets and a ring of three chunks.
template <int K, bool FMA, bool WDEV, int CH, int NS, int ND, bool EDGE>
__device__ __forceinline__ void ring_body(const LocalArgs& a, float4* __restrict__ lds,
                                          const float* __restrict__ in, float* __restrict__ out,
                                          int xl, int y0, int nout) {
  constexpr int V = kRingV, QR = CH / 4;
  constexpr int R = K / 2;
  constexpr int NT = QR + K - 1;           // LDS rows a wave reads per chunk
  static_assert(R <= V && 2 * R <= CH, "mask radius exceeds the ring's halo");
  // Split edge loads (16-byte + per-pixel) hold twice the staged registers;
  // the larger masks, whose VALU work hides the per-pixel loads, keep the
  // registers and load every pixel of an edge strip separately.
  constexpr bool SPLIT = EDGE && K <= 5;
  constexpr int NSLOT = EDGE ? 3 : ND + 2;   // chunks in the LDS ring
  const Geom& g = a.g;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane((threadIdx.x >> 6) & 3);
  const int nin = nout + K - 1;            // input rows of the segment
  const int iy0 = g.in_oy + y0 - R;        // absolute row of input row 0
  const int ix = g.in_ox + xl;             // absolute column of the lane's pixel 0
  const bool cmode = g.boundary == HIPACC_BOUNDARY_CONSTANT;
  const float cv = g.const_f;
  // edge strips: per-lane byte offsets of the 16-byte load (raw lanes) and
  // of the per-pixel loads (boundary-mapped columns; CONSTANT pixels read
  // nothing and are selected at the LDS write)
  int roff = kRingDrop;
  int eoff[V];
  bool ecst[V];
  if constexpr (EDGE) {
    int lx, hx, ly, hy;
    raw_region(g, lx, hx, ly, hy);
    const bool lane_raw = SPLIT && a.in_vec_ok && ix >= lx && ix + V <= hx;
    // lanes farther than R columns from every stored column feed no output
    // (the iteration space ends inside the strip): they load nothing
    const int s0 = xl - lane * V + V, s1 = min(s0 + kRingStripW, g.is_w);
    const bool used = xl + V > s0 - R && xl < s1 + R;
    const LaneCols<V> cols = lane_cols<V>(g, ix);
    roff = lane_raw && used ? ix * 4 : kRingDrop;
#pragma unroll
    for (int e = 0; e < V; ++e) {
      ecst[e] = !lane_raw && cols.cst[e];
      eoff[e] = (lane_raw || cols.cst[e] || !used) ? kRingDrop : cols.x[e] * 4;
    }
  }
  const int in_row_bytes = g.in_imgw * 4;

  auto load_chunk = [&](int c) {
    RingStage<SPLIT, QR> st;
    st.rc = 0;
#pragma unroll
    for (int i = 0; i < QR; ++i) {
      const int k = min(c * CH + wave + 4 * i, nin - 1);
      bool rc;
      const int y = row_src(g, iy0 + k, rc);
      st.rc |= rc ? 1 << i : 0;
      const float* row = in + (long long)y * g.in_stride;
      if constexpr (!EDGE) {
        st.r[i] = *reinterpret_cast<const float4*>(row + ix);
      } else {
        const __amdgpu_buffer_rsrc_t rs = ring_rsrc(row, in_row_bytes);
        float v[V];
#pragma unroll
        for (int e = 0; e < V; ++e)
          v[e] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, eoff[e], 0, 0));
        if constexpr (SPLIT) {
          st.r[i] = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(rs, roff, 0, 0));
          st.p[i] = make_float4(v[0], v[1], v[2], v[3]);
        } else {
          st.r[i] = make_float4(v[0], v[1], v[2], v[3]);
        }
      }
    }
    return st;
  };
  // CONSTANT pixels are selected here, after the loads have landed.
  auto write_chunk = [&](int c, const RingStage<SPLIT, QR>& st) {
    const int slot = c % NSLOT;
#pragma unroll
    for (int i = 0; i < QR; ++i) {
      float4 v;
      if constexpr (SPLIT) v = ring_merge(st.r[i], st.p[i]);
      else v = st.r[i];
      if (cmode) {
        const bool rc = (st.rc >> i) & 1;
        if constexpr (EDGE) {
          v.x = (rc || ecst[0]) ? cv : v.x;
          v.y = (rc || ecst[1]) ? cv : v.y;
          v.z = (rc || ecst[2]) ? cv : v.z;
          v.w = (rc || ecst[3]) ? cv : v.w;
        } else {
          v = rc ? make_float4(cv, cv, cv, cv) : v;
        }
      }
      lds[(slot * CH + wave + 4 * i) * 64 + lane] = v;
    }
  };

  // stores: one 16-byte store per lane whose 4 pixels are all inside the
  // iteration space (interior strips: every stored lane), else one 4-byte
  // store per pixel inside it
  const bool store_lane = lane >= 1 && lane < 63;
  const bool full = store_lane && (!EDGE || (a.out_vec_ok && xl + V <= g.is_w));
  const int voff = full ? xl * 4 : kRingDrop;
  int soff[V];
  if constexpr (EDGE) {
#pragma unroll
    for (int e = 0; e < V; ++e)
      soff[e] = (store_lane && !full && xl + e < g.is_w) ? (xl + e) * 4 : kRingDrop;
  }
  float* obase = out + (long long)(g.out_oy + y0) * g.out_stride + g.out_ox;
  const int out_row_bytes = g.is_w * 4;

  auto compute = [&](int s) {
    const int jw = wave * QR;              // the wave's first output row in the chunk
    const int j0 = s * CH + jw;
    float acc[QR][V];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const int off = jw + t;              // row of chunk s (or of chunk s+1 past CH)
      const int c = off >= CH ? 1 : 0;
      const int slot = (s + c) % NSLOT;
      const float4* rp = lds + (slot * CH + off - c * CH) * 64 + lane;
      const float4 q = *rp;
      constexpr int NWIN = V + 2 * R;
      float win[NWIN];
      win[R] = q.x;
      win[R + 1] = q.y;
      win[R + 2] = q.z;
      win[R + 3] = q.w;
      if constexpr (R == V) {
        // 9x9: the neighbour lanes' whole words, read from LDS at fixed
        // offsets of the same address (8 VALU lane shifts fewer per row;
        // lanes 0 and 63 read a guard word or a neighbouring row, and what
        // they compute is never stored).  Narrower reads of a neighbour's
        // word conflict on the LDS banks, so smaller masks shift lanes.
        const float4 p = rp[-1], n = rp[1];
        win[0] = p.x;
        win[1] = p.y;
        win[2] = p.z;
        win[3] = p.w;
        win[R + V] = n.x;
        win[R + V + 1] = n.y;
        win[R + V + 2] = n.z;
        win[R + V + 3] = n.w;
      } else {
        const float cur[V] = {q.x, q.y, q.z, q.w};
#pragma unroll
        for (int k = 0; k < R; ++k) {
          win[k] = from_prev_lane(cur[V - R + k]);
          win[R + V + k] = from_next_lane(cur[k]);
        }
      }
#pragma unroll
      for (int qq = 0; qq < QR; ++qq) {
        const int dyi = t - qq;
        if (dyi < 0 || dyi >= K) continue;
#pragma unroll
        for (int dx = 0; dx < K; ++dx) {
          const float w = StreamTaps<LocalArgs>::template w2d<float, WDEV>(a, dyi * K + dx);
#pragma unroll
          for (int e = 0; e < V; ++e) {
            if (dyi == 0 && dx == 0) acc[qq][e] = mul_first(w, win[e]);
            else acc[qq][e] = madd<FMA>(w, win[e + dx], acc[qq][e]);
          }
        }
      }
    }
#pragma unroll
    for (int qq = 0; qq < QR; ++qq) {
      const int j = j0 + qq;
      const __amdgpu_buffer_rsrc_t rs =
          ring_rsrc(obase + (long long)j * g.out_stride, j < nout ? out_row_bytes : 0);
      const ring_u4 v = {__builtin_bit_cast(unsigned, acc[qq][0]),
                         __builtin_bit_cast(unsigned, acc[qq][1]),
                         __builtin_bit_cast(unsigned, acc[qq][2]),
                         __builtin_bit_cast(unsigned, acc[qq][3])};
      __builtin_amdgcn_raw_buffer_store_b128(v, rs, voff, 0, 2);   // 2: nontemporal
      if constexpr (EDGE) {
#pragma unroll
        for (int e = 0; e < V; ++e)
          __builtin_amdgcn_raw_buffer_store_b32(v[e], rs, soff[e], 0, 2);
      }
    }
  };

  const int nos = (nout + CH - 1) / CH;
  if constexpr (!EDGE) {
    // Direct loads: chunk c lands in slot c % NSLOT.  Step s waits for its
    // own loads of chunk s+1, meets the other waves at the barrier, issues
    // the loads of chunk s+ND+1 into the slot chunk s-1 held, and computes
    // output chunk s from chunks s and s+1.  The wait leaves at most the
    // QR*(ND-1) loads issued after chunk s+1 outstanding: loads retire in
    // issue order, so while a load of chunk s+1 is outstanding every later
    // load is too and the count stays above that bound, whatever stores
    // are pending in the same counter.
    auto issue = [&](int c) {
      const int slot = c % NSLOT;
#pragma unroll
      for (int i = 0; i < QR; ++i) {
        const int k = min(c * CH + wave + 4 * i, nin - 1);
        bool rc;
        const int y = row_src(g, iy0 + k, rc);
        __builtin_amdgcn_global_load_lds(
            static_cast<const void*>(in + (long long)y * g.in_stride + ix),
            (__attribute__((address_space(3))) void*)(lds + (slot * CH + wave + 4 * i) * 64), 16,
            0, 0);
      }
    };
#pragma unroll
    for (int c = 0; c <= ND; ++c) issue(c);
    for (int s = 0; s < nos; ++s) {
      ring_wait_vm<QR * (ND - 1)>();
      ring_barrier();
      issue(s + ND + 1);
      compute(s);
    }
    ring_wait_vm<0>();                   // no load may land after the block ends
  } else {
    // Register sets: chunk c lives in LDS slot c % 3.  Step s computes output
    // chunk s from chunks s and s+1, after writing chunk s+2 (loaded NS steps
    // earlier) and issuing the loads of chunk s+2+NS.
    RingStage<SPLIT, QR> sets[NS];
    {
      const RingStage<SPLIT, QR> c0 = load_chunk(0), c1 = load_chunk(1);
      write_chunk(0, c0);
      write_chunk(1, c1);
    }
#pragma unroll
    for (int u = 0; u < NS; ++u) sets[(u + 2) % NS] = load_chunk(u + 2);
    __syncthreads();
    int s = 0;
    for (; s + NS <= nos; s += NS) {
#pragma unroll
      for (int u = 0; u < NS; ++u) {
        const int k = (u + 2) % NS;
        write_chunk(s + u + 2, sets[k]);
        sets[k] = load_chunk(s + u + 2 + NS);
        compute(s + u);
        __syncthreads();
      }
    }
#pragma unroll
    for (int u = 0; u < NS; ++u) {
      if (s + u < nos) {
        const int k = (u + 2) % NS;
        write_chunk(s + u + 2, sets[k]);
        compute(s + u);
        __syncthreads();
      }
    }
  }
}

// Item i of the launch (one segment of one strip of one image).
struct RingItem {
  int b, strip, y0, nout;
};
__device__ __forceinline__ RingItem ring_item(const RingGeom& rg, const Geom& g, int item) {
  RingItem it;
  it.b = item / rg.per_img;
  int t = item - it.b * rg.per_img;
  int rows;
  const int nm = rg.strips - rg.nl - rg.nr;
  if (t < rg.nl * rg.segs_e) {
    it.strip = t / rg.segs_e;
    t -= it.strip * rg.segs_e;
    rows = rg.seg_rows_e;
  } else if ((t -= rg.nl * rg.segs_e) < nm * rg.segs) {
    it.strip = t / rg.segs;
    t -= it.strip * rg.segs;
    it.strip += rg.nl;
    rows = rg.seg_rows;
  } else {
    t -= nm * rg.segs;
    it.strip = t / rg.segs_e;
    t -= it.strip * rg.segs_e;
    it.strip += rg.nl + nm;
    rows = rg.seg_rows_e;
  }
  it.y0 = t * rows;
  it.nout = min(rows, g.is_h - it.y0);
  return it;
}

// One item per block of 4 waves.
template <int K, bool FMA, bool WDEV, int CH, int NS, int ND>
__global__ __launch_bounds__(256) void ring_kernel(const LocalArgs a, const RingGeom rg) {
  // the ring with one guard word on each side
  __shared__ __align__(16) float4 lds_ring[(ND + 2 > 3 ? ND + 2 : 3) * CH * 64 + 2];
  float4* lds = lds_ring + 1;
  const Geom& g = a.g;
  if ((int)blockIdx.x >= rg.items) return;
  const RingItem it = ring_item(rg, g, blockIdx.x);
  const float* in = static_cast<const float*>(g.in) + (long long)it.b * g.in_plane;
  float* out = static_cast<float*>(g.out) + (long long)it.b * g.out_plane;
  const int lane = threadIdx.x & 63;
  const int xs = it.strip * kRingStripW - kRingV;   // iteration-space column of lane 0's pixel 0
  int lx, hx, ly, hy;
  raw_region(g, lx, hx, ly, hy);
  const int ix0 = g.in_ox + xs;
  // interior: raw columns, full vector stores, and (CONSTANT) no row outside
  // the window -- the direct loads cannot substitute the constant
  const int iy0 = g.in_oy + it.y0 - K / 2;
  const bool raw = a.in_vec_ok && a.out_vec_ok && ix0 >= lx && ix0 + 64 * kRingV <= hx &&
                   xs + 63 * kRingV <= g.is_w &&
                   (g.boundary != HIPACC_BOUNDARY_CONSTANT ||
                    (iy0 >= g.lo_y && iy0 + it.nout + K - 1 <= g.hi_y));
  if (raw)
    ring_body<K, FMA, WDEV, CH, NS, ND, false>(a, lds, in, out, xs + lane * kRingV, it.y0,
                                               it.nout);
  else
    ring_body<K, FMA, WDEV, CH, NS, ND, true>(a, lds, in, out, xs + lane * kRingV, it.y0,
                                              it.nout);
}

// ---------------------------------------------------------------------------
// Separable masks (f32 -> f32, u8 -> u8).  The strip keeps HL halo lanes on each side
// (HL = ceil(RX / 4)), so 64 - 2 HL lanes are stored.  Each wave stages CH/4
// input rows of a chunk in registers (NS sets in flight), computes their
// horizontal pass (taps from the neighbour lanes through DPP) and writes the
// intermediate rows into a 3-chunk LDS ring; after the barrier each wave
// sums KY intermediate rows per output row.  The arithmetic equals the
// streaming separable kernel's (hh_stream.hpp) operation for operation: the
// horizontal sum in tap order, the intermediate row replaced by the
// constant for a row outside the window in CONSTANT mode, then the vertical
// sum in tap order, each sum starting with a plain product.
// ---------------------------------------------------------------------------
// Staged input rows of the separable ring: 4 pixels per lane and row.
template <typename TI, int QR> struct RingSepStage {
  float4 r[QR];
  int rc;
};
template <int QR> struct RingSepStage<uint8_t, QR> {
  unsigned r[QR];   // 4 packed pixels
  int rc;
};

// ISUM: integer weights and pixels, summed exactly in float (the host admits
// only masks whose partial sums stay below 2^24), integer output.
template <typename TI, typename TO, int KX, int KY, bool FMA, bool ISUM, bool EDGE, int CH, int NS>
__device__ __forceinline__ void ring_sep_body(const SepArgs& a, float4* __restrict__ lds,
                                              const TI* __restrict__ in, TO* __restrict__ out,
                                              int xl, int y0, int nout) {
  constexpr int V = kRingV, QR = CH / 4;
  constexpr int RX = KX / 2, RY = KY / 2;
  constexpr int HL = (RX + V - 1) / V;
  constexpr int NT = QR + KY - 1;
  constexpr int NWIN = V + 2 * RX;
  constexpr bool U8 = std::is_same<TI, uint8_t>::value;
  constexpr int BI = (int)sizeof(TI), BO = (int)sizeof(TO);
  static_assert(2 * RY <= CH, "vertical radius exceeds the ring's halo");
  const Geom& g = a.g;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int nin = nout + KY - 1;
  const int iy0 = g.in_oy + y0 - RY;
  const int ix = g.in_ox + xl;
  const bool cmode = g.boundary == HIPACC_BOUNDARY_CONSTANT;
  const float cpx = (float)const_pixel<TI>(g);   // a pixel outside the window
  // an intermediate row outside the window
  const float crow = ISUM ? (float)g.const_i : g.const_f;
  float wr[KX], wc[KY];
#pragma unroll
  for (int i = 0; i < KX; ++i) wr[i] = ISUM ? (float)a.rowi[i] : a.rowf[i];
#pragma unroll
  for (int i = 0; i < KY; ++i) wc[i] = ISUM ? (float)a.coli[i] : a.colf[i];
  const float rnd = g.cvt == HIPACC_CVT_ROUND ? 0.5f : 0.0f;
  int eoff[V];
  bool ecst[V];
  if constexpr (EDGE) {
    const LaneCols<V> cols = lane_cols<V>(g, ix);
#pragma unroll
    for (int e = 0; e < V; ++e) {
      ecst[e] = cols.cst[e];
      eoff[e] = cols.cst[e] ? kRingDrop : cols.x[e] * BI;
    }
  }
  const int in_row_bytes = g.in_imgw * BI;

  auto load_chunk = [&](int c) {
    RingSepStage<TI, QR> st;
    st.rc = 0;
#pragma unroll
    for (int i = 0; i < QR; ++i) {
      const int k = min(c * CH + wave + 4 * i, nin - 1);
      bool rc;
      const int y = row_src(g, iy0 + k, rc);
      st.rc |= rc ? 1 << i : 0;
      const TI* row = in + (long long)y * g.in_stride;
      if constexpr (!EDGE) {
        if constexpr (U8) st.r[i] = *reinterpret_cast<const unsigned*>(row + ix);
        else st.r[i] = *reinterpret_cast<const float4*>(row + ix);
      } else {
        const __amdgpu_buffer_rsrc_t rs = ring_rsrc(row, in_row_bytes);
        if constexpr (U8) {
          unsigned w = 0;
#pragma unroll
          for (int e = 0; e < V; ++e)
            w |= (unsigned)__builtin_amdgcn_raw_buffer_load_b8(rs, eoff[e], 0, 0) << (8 * e);
          st.r[i] = w;
        } else {
          float v[V];
#pragma unroll
          for (int e = 0; e < V; ++e)
            v[e] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, eoff[e], 0, 0));
          st.r[i] = make_float4(v[0], v[1], v[2], v[3]);
        }
      }
    }
    return st;
  };
  // horizontal pass of the staged rows, written to LDS slot c % 3
  auto write_chunk = [&](int c, const RingSepStage<TI, QR>& st) {
    const int slot = c % 3;
#pragma unroll
    for (int i = 0; i < QR; ++i) {
      float cur[V];
      if constexpr (U8) {
#pragma unroll
        for (int e = 0; e < V; ++e) cur[e] = (float)((st.r[i] >> (8 * e)) & 0xffu);
      } else {
        cur[0] = st.r[i].x;
        cur[1] = st.r[i].y;
        cur[2] = st.r[i].z;
        cur[3] = st.r[i].w;
      }
      if constexpr (EDGE) {
        if (cmode) {
#pragma unroll
          for (int e = 0; e < V; ++e) cur[e] = ecst[e] ? cpx : cur[e];
        }
      }
      float win[NWIN];
#pragma unroll
      for (int wi = 0; wi < NWIN; ++wi) {
        const int rel = wi - RX;
        const int d = rel >= 0 ? rel / V : -((-rel + V - 1) / V);
        const int e = rel - d * V;
        win[wi] = d == 0 ? cur[e] : lane_shift(cur[e], d);
      }
      float h[V];
#pragma unroll
      for (int e = 0; e < V; ++e) {
        float s0 = 0.0f;
#pragma unroll
        for (int dx = 0; dx < KX; ++dx)
          s0 = dx == 0 ? mul_first(wr[dx], win[e]) : madd<FMA>(wr[dx], win[e + dx], s0);
        h[e] = s0;
      }
      if (cmode && ((st.rc >> i) & 1)) {
#pragma unroll
        for (int e = 0; e < V; ++e) h[e] = crow;
      }
      lds[(slot * CH + wave + 4 * i) * 64 + lane] = make_float4(h[0], h[1], h[2], h[3]);
    }
  };

  const bool store_lane = lane >= HL && lane < 64 - HL;
  const bool full = store_lane && (!EDGE || (a.out_vec_ok && xl + V <= g.is_w));
  const int voff = full ? xl * BO : kRingDrop;
  int soff[V];
  if constexpr (EDGE) {
#pragma unroll
    for (int e = 0; e < V; ++e)
      soff[e] = (store_lane && !full && xl + e < g.is_w) ? (xl + e) * BO : kRingDrop;
  }
  TO* obase = out + (long long)(g.out_oy + y0) * g.out_stride + g.out_ox;
  const int out_row_bytes = g.is_w * BO;

  auto compute = [&](int s) {
    const int jw = wave * QR;
    const int j0 = s * CH + jw;
    float acc[QR][V];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const int off = jw + t;
      const int c = off >= CH ? 1 : 0;
      const int slot = (s + c) % 3;
      const float4 q = lds[(slot * CH + off - c * CH) * 64 + lane];
      const float hv[V] = {q.x, q.y, q.z, q.w};
#pragma unroll
      for (int qq = 0; qq < QR; ++qq) {
        const int dyi = t - qq;
        if (dyi < 0 || dyi >= KY) continue;
        const float w = wc[dyi];
#pragma unroll
        for (int e = 0; e < V; ++e)
          acc[qq][e] = dyi == 0 ? mul_first(w, hv[e]) : madd<FMA>(w, hv[e], acc[qq][e]);
      }
    }
#pragma unroll
    for (int qq = 0; qq < QR; ++qq) {
      const int j = j0 + qq;
      const __amdgpu_buffer_rsrc_t rs =
          ring_rsrc(obase + (long long)j * g.out_stride, j < nout ? out_row_bytes : 0);
      if constexpr (std::is_same<TO, float>::value) {
        const ring_u4 v = {__builtin_bit_cast(unsigned, acc[qq][0]),
                           __builtin_bit_cast(unsigned, acc[qq][1]),
                           __builtin_bit_cast(unsigned, acc[qq][2]),
                           __builtin_bit_cast(unsigned, acc[qq][3])};
        __builtin_amdgcn_raw_buffer_store_b128(v, rs, voff, 0, 2);
        if constexpr (EDGE) {
#pragma unroll
          for (int e = 0; e < V; ++e) __builtin_amdgcn_raw_buffer_store_b32(v[e], rs, soff[e], 0, 2);
        }
      } else {
        TO px[V];
#pragma unroll
        for (int e = 0; e < V; ++e)
          px[e] = ISUM ? cvt_out<TO>::from_i((int32_t)acc[qq][e])
                       : cvt_out<TO>::from_f(acc[qq][e] + rnd);
        if constexpr (BO == 1) {
          unsigned w = 0;
#pragma unroll
          for (int e = 0; e < V; ++e) w |= (unsigned)(uint8_t)px[e] << (8 * e);
          __builtin_amdgcn_raw_buffer_store_b32(w, rs, voff, 0, 2);
          if constexpr (EDGE) {
#pragma unroll
            for (int e = 0; e < V; ++e)
              __builtin_amdgcn_raw_buffer_store_b8((unsigned char)px[e], rs, soff[e], 0, 2);
          }
        } else if constexpr (BO == 2) {
          typedef unsigned int ring_u2 __attribute__((ext_vector_type(2)));
          const ring_u2 v = {(unsigned)(uint16_t)px[0] | (unsigned)(uint16_t)px[1] << 16,
                             (unsigned)(uint16_t)px[2] | (unsigned)(uint16_t)px[3] << 16};
          __builtin_amdgcn_raw_buffer_store_b64(v, rs, voff, 0, 2);
          if constexpr (EDGE) {
#pragma unroll
            for (int e = 0; e < V; ++e)
              __builtin_amdgcn_raw_buffer_store_b16((unsigned short)px[e], rs, soff[e], 0, 2);
          }
        } else {
          const ring_u4 v = {__builtin_bit_cast(unsigned, px[0]), __builtin_bit_cast(unsigned, px[1]),
                             __builtin_bit_cast(unsigned, px[2]), __builtin_bit_cast(unsigned, px[3])};
          __builtin_amdgcn_raw_buffer_store_b128(v, rs, voff, 0, 2);
          if constexpr (EDGE) {
#pragma unroll
            for (int e = 0; e < V; ++e)
              __builtin_amdgcn_raw_buffer_store_b32(v[e], rs, soff[e], 0, 2);
          }
        }
      }
    }
  };

  const int nos = (nout + CH - 1) / CH;
  RingSepStage<TI, QR> sets[NS];
  {
    const RingSepStage<TI, QR> c0 = load_chunk(0), c1 = load_chunk(1);
    write_chunk(0, c0);
    write_chunk(1, c1);
  }
#pragma unroll
  for (int u = 0; u < NS; ++u) sets[(u + 2) % NS] = load_chunk(u + 2);
  __syncthreads();
  int s = 0;
  for (; s + NS <= nos; s += NS) {
#pragma unroll
    for (int u = 0; u < NS; ++u) {
      const int k = (u + 2) % NS;
      write_chunk(s + u + 2, sets[k]);
      sets[k] = load_chunk(s + u + 2 + NS);
      compute(s + u);
      __syncthreads();
    }
  }
#pragma unroll
  for (int u = 0; u < NS; ++u) {
    if (s + u < nos) {
      const int k = (u + 2) % NS;
      write_chunk(s + u + 2, sets[k]);
      compute(s + u);
      __syncthreads();
    }
  }
}

template <typename TI, typename TO, int KX, int KY, bool FMA, bool ISUM, int CH, int NS>
__global__ __launch_bounds__(256) void ring_sep_kernel(const SepArgs a, const RingGeom rg) {
  constexpr int HL = (KX / 2 + kRingV - 1) / kRingV;
  constexpr int SW = (64 - 2 * HL) * kRingV;
  __shared__ __align__(16) float4 lds[3 * CH * 64];
  const Geom& g = a.g;
  if ((int)blockIdx.x >= rg.items) return;
  const RingItem it = ring_item(rg, g, blockIdx.x);
  const TI* in = static_cast<const TI*>(g.in) + (long long)it.b * g.in_plane;
  TO* out = static_cast<TO*>(g.out) + (long long)it.b * g.out_plane;
  const int lane = threadIdx.x & 63;
  const int xs = it.strip * SW - HL * kRingV;
  int lx, hx, ly, hy;
  raw_region(g, lx, hx, ly, hy);
  const int ix0 = g.in_ox + xs;
  const bool raw = a.in_vec_ok && a.out_vec_ok && ix0 >= lx && ix0 + 64 * kRingV <= hx &&
                   xs + (64 - HL) * kRingV <= g.is_w;
  if (raw)
    ring_sep_body<TI, TO, KX, KY, FMA, ISUM, false, CH, NS>(a, lds, in, out, xs + lane * kRingV,
                                                            it.y0, it.nout);
  else
    ring_sep_body<TI, TO, KX, KY, FMA, ISUM, true, CH, NS>(a, lds, in, out, xs + lane * kRingV,
                                                           it.y0, it.nout);
}

}  // namespace hh
