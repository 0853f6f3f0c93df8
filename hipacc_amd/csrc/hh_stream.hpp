// hh_stream.hpp -- streaming local operators (weighted SUM 2D masks,
// separable masks, box sums) for the HBM-bound hot path.
//
// Each wave owns a vertical strip of the iteration space: lane l holds 4
// adjacent columns, so one row of the strip is one aligned 4-pixel load per
// lane.  The wave walks its segment of rows top to bottom:
//  - rows are prefetched PF rows ahead into a register ring, so every lane
//    keeps PF loads in flight without LDS or barriers;
//  - the horizontal halo comes from neighbouring lanes by cross-lane shuffles;
//    the HL lanes at either edge only supply halo (they load, but store
//    nothing), so adjacent strips overlap by 2*HL*4 columns instead of issuing
//    separate halo loads;
//  - every loaded row is added into the KY rolling accumulators of the output
//    rows whose window contains it, so each output pixel receives its taps in
//    row-major order (dy outer, dx inner), the order of lib/AST/
//    Convolution.cpp:397-435 and of the DSL's ElementIterator;
//  - boundary handling is a scalar remap of the row index (or a constant
//    row) plus a per-element column remap only for lanes whose 4 columns cross
//    the accessor window (lib/AST/BorderHandling.cpp:41-120).
// Consecutive waves take vertically adjacent segments of one strip, and the
// block order is permuted so that blocks sharing an XCD (blockIdx % 8) take
// consecutive tiles: the halo rows two segments share hit the same L2.
#pragma once

#include "hh_tile.hpp"

namespace hh {

constexpr int kStreamVec = 4;  // pixels per lane

enum StreamKind { kStream2D = 0, kStreamSep = 1, kStreamBox = 2 };

struct StreamGeom {
  int strips, segs, seg_rows;  // tiling of one image
  int strip_w;                 // useful columns per strip
  int total_waves;
  int xcd_perm;                // permute blocks across the 8 XCDs
};

template <typename T> __device__ __forceinline__ T shfl_up1(T v, int d) { return __shfl_up(v, d, 64); }
template <typename T> __device__ __forceinline__ T shfl_dn1(T v, int d) { return __shfl_down(v, d, 64); }

// Loads the 4 pixels of one lane for input row iy (absolute), applying the
// boundary mode.
template <typename TI>
__device__ __forceinline__ void stream_load(TI (&v)[kStreamVec], const TI* __restrict__ in,
                                            const Geom& g, int ix, int iy, bool vec_ok, int lx,
                                            int hx) {
  const bool cmode = g.boundary == HIPACC_BOUNDARY_CONSTANT;
  if (cmode && (iy < g.lo_y || iy >= g.hi_y)) {
    const TI c = const_pixel<TI>(g);
#pragma unroll
    for (int e = 0; e < kStreamVec; ++e) v[e] = c;
    return;
  }
  if (!cmode) iy = bh_map(g.boundary, iy, g.lo_y, g.hi_y, g.in_imgh);
  const TI* row = in + (long long)iy * g.in_stride;
  if (vec_ok && ix >= lx && ix + kStreamVec <= hx) {
    using W = typename word_of<kStreamVec * sizeof(TI)>::type;
    const W w = *reinterpret_cast<const W*>(row + ix);
    __builtin_memcpy(v, &w, sizeof(W));
  } else {
#pragma unroll
    for (int e = 0; e < kStreamVec; ++e) {
      int x = ix + e;
      if (cmode && (x < g.lo_x || x >= g.hi_x)) {
        v[e] = const_pixel<TI>(g);
      } else {
        if (!cmode) x = bh_map(g.boundary, x, g.lo_x, g.hi_x, g.in_imgw);
        v[e] = row[x];
      }
    }
  }
}

template <typename ArgsT> struct StreamTaps;
template <> struct StreamTaps<LocalArgs> {
  template <typename ACC, bool WDEV>
  __device__ __forceinline__ static ACC w2d(const LocalArgs& a, int t) {
    if constexpr (WDEV) return (ACC)a.wdev[t];
    else if constexpr (std::is_same<ACC, float>::value) return a.w.f[t];
    else return (ACC)a.w.i[t];
  }
};

template <typename TI, typename TO, typename ACC, int KX, int KY, int PF, bool FMA, int KIND,
          bool WDEV, typename ArgsT>
__global__ __launch_bounds__(256) void stream_kernel(const ArgsT a, const StreamGeom sg) {
  constexpr int V = kStreamVec;
  constexpr int RX = KX / 2, RY = KY / 2;
  constexpr int HL = (RX + V - 1) / V;          // halo lanes on each side
  constexpr int NWIN = V + 2 * RX;
  constexpr bool IS_F = std::is_same<ACC, float>::value;
  static_assert(PF % KY == 0, "prefetch ring must be a multiple of KY");

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

  const TI* in = static_cast<const TI*>(g.in) + (long long)b * g.in_plane;
  TO* out = static_cast<TO*>(g.out) + (long long)b * g.out_plane;
  const int xl = strip * sg.strip_w - HL * V + lane * V;  // iteration-space column of element 0
  const int ix = g.in_ox + xl;
  const int y0 = seg * sg.seg_rows;
  const int nout = min(sg.seg_rows, g.is_h - y0);
  const int nj = nout + KY - 1;                   // input rows streamed
  const int iy0 = g.in_oy + y0 - RY;
  int lx, hx, ly, hy;
  raw_region(g, lx, hx, ly, hy);
  const bool vec_ok = a.in_vec_ok;
  const bool store_lane = lane >= HL && lane < 64 - HL;
  const bool cmode = g.boundary == HIPACC_BOUNDARY_CONSTANT;

  TI ring[PF][V];
  ACC acc[KY][V];
#pragma unroll
  for (int p = 0; p < PF; ++p)
    if (p < nj) stream_load<TI>(ring[p], in, g, ix, iy0 + p, vec_ok, lx, hx);

  for (int jb = 0; jb < nj; jb += PF) {
#pragma unroll
    for (int u = 0; u < PF; ++u) {
      const int j = jb + u;
      if (j >= nj) break;
      ACC cur[V];
#pragma unroll
      for (int e = 0; e < V; ++e) cur[e] = (ACC)ring[u][e];
      if (j + PF < nj) stream_load<TI>(ring[u], in, g, ix, iy0 + j + PF, vec_ok, lx, hx);

      // window of this lane's row: columns xl-RX .. xl+V-1+RX
      ACC win[NWIN];
#pragma unroll
      for (int wi = 0; wi < NWIN; ++wi) {
        const int rel = wi - RX;                          // column relative to element 0
        const int d = rel >= 0 ? rel / V : -((-rel + V - 1) / V);  // source lane offset
        const int e = rel - d * V;
        if (d == 0) win[wi] = cur[e];
        else if (d < 0) win[wi] = shfl_up1(cur[e], -d);
        else win[wi] = shfl_dn1(cur[e], d);
      }

      const int slot0 = u % KY;                           // accumulator of output row j
      if constexpr (KIND == kStream2D) {
#pragma unroll
        for (int dyi = 0; dyi < KY; ++dyi) {
          const int s = (slot0 - dyi + KY) % KY;          // output row j - dyi
          if (dyi == 0) {
#pragma unroll
            for (int e = 0; e < V; ++e) acc[s][e] = ACC(0);
          }
#pragma unroll
          for (int dx = 0; dx < KX; ++dx) {
            const ACC w = StreamTaps<ArgsT>::template w2d<ACC, WDEV>(a, dyi * KX + dx);
#pragma unroll
            for (int e = 0; e < V; ++e) acc[s][e] = madd<FMA>(w, win[e + dx], acc[s][e]);
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
              s0 = madd<FMA>(w, win[e + dx], s0);
            }
            h[e] = s0;
          }
        }
        if (cmode) {
          const int iy = iy0 + j;
          if (iy < g.lo_y || iy >= g.hi_y) {
#pragma unroll
            for (int e = 0; e < V; ++e) h[e] = IS_F ? (ACC)g.const_f : (ACC)g.const_i;
          }
        }
#pragma unroll
        for (int dyi = 0; dyi < KY; ++dyi) {
          const int s = (slot0 - dyi + KY) % KY;
          if (dyi == 0) {
#pragma unroll
            for (int e = 0; e < V; ++e) acc[s][e] = ACC(0);
          }
          if constexpr (KIND == kStreamBox) {
#pragma unroll
            for (int e = 0; e < V; ++e) acc[s][e] = (ACC)((uint32_t)acc[s][e] + (uint32_t)h[e]);
          } else {
            const ACC w = IS_F ? (ACC)a.colf[dyi] : (ACC)a.coli[dyi];
#pragma unroll
            for (int e = 0; e < V; ++e) acc[s][e] = madd<FMA>(w, h[e], acc[s][e]);
          }
        }
      }
      // output row j-(KY-1) is complete
      const int o = j - (KY - 1);
      if (o >= 0 && store_lane) {
        const int s = (slot0 + 1) % KY;  // == (slot0 - (KY-1)) mod KY
        store_row<TO, ACC, V>(out, g, a.out_vec_ok, xl, y0 + o, acc[s]);
      }
    }
  }
}

}  // namespace hh
