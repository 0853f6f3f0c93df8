// hh_ring8.hpp -- the block row ring of hh_ring.hpp for 8-bit input pixels:
// u8 -> u8 with a float sum (rounded and saturated as requested) and
// u8 -> i16 / u32 / i32 with an integer sum, square 2D masks 3x3 and 5x5.
//
// A lane holds 16 adjacent pixels (one 16-byte word), so one wave row
// stages 1024 columns.  Lane 0 and lane sl+1 only supply the horizontal
// halo; the sl lanes between them are stored.  The host picks sl so that the
// strips of a row are equally wide (the lanes past sl+1 idle) rather than
// leaving a narrow last strip.  The LDS ring, the direct loads of interior
// strips, the register sets of edge strips and the wait discipline are those
// of ring_body; the 16 bytes are widened to the sum type when a row is read
// from LDS, and the neighbour pixels come from one DPP move of the
// neighbour's first or last word.
//
// The taps are applied in row-major order with a plain product first
// (lib/AST/Convolution.cpp:397-435), as in the streaming kernel of
// hh_stream.hpp, so the float sums equal its sums bit for bit.  Integer
// sums (ISUM) are also formed in float: the host admits only masks with
// 255 * sum|w| < 2^24, so every partial sum is an integer that float
// represents exactly and the result equals the DSL's int arithmetic, at the
// full-rate FMA instead of a 32-bit integer multiply.  Boundary modes follow
// lib/AST/BorderHandling.cpp:41-120 exactly as in ring_body.
#pragma once

#include "hh_ring.hpp"

namespace hh {

constexpr int kRing8V = 16;   // pixels per lane

// Staged rows of an edge strip: the 16-byte load of the lanes inside the
// raw-readable columns and the per-pixel loads of the other lanes, one
// register per pixel (the load that does not apply to a lane is dropped and
// reads 0).
template <int QR> struct Ring8Stage {
  uint4 r[QR];
  unsigned p[QR][kRing8V];
  int rc;          // bit i: staged row i is outside the window in CONSTANT mode
};

// Byte k of w as a float.
__device__ __forceinline__ float ring8_px(unsigned w, int k) {
  return (float)((w >> (8 * k)) & 0xffu);
}

// Output words of a lane: 16 pixels of TO packed into 16 * sizeof(TO) / 4
// dwords.  ISUM: the sum is an exact integer.
template <typename TO, bool ISUM>
__device__ __forceinline__ void ring8_pack(const float (&acc)[kRing8V], float rnd, TO (&px)[kRing8V],
                                           unsigned (&d)[kRing8V * sizeof(TO) / 4]) {
  constexpr int BO = (int)sizeof(TO);
#pragma unroll
  for (int e = 0; e < kRing8V; ++e) {
    if constexpr (ISUM) px[e] = cvt_out<TO>::from_i((int32_t)acc[e]);
    else px[e] = cvt_out<TO>::from_f(acc[e] + rnd);
  }
#pragma unroll
  for (int k = 0; k < kRing8V * BO / 4; ++k) {
    if constexpr (BO == 1) {
      d[k] = (unsigned)(uint8_t)px[4 * k] | (unsigned)(uint8_t)px[4 * k + 1] << 8 |
             (unsigned)(uint8_t)px[4 * k + 2] << 16 | (unsigned)(uint8_t)px[4 * k + 3] << 24;
    } else if constexpr (BO == 2) {
      d[k] = (unsigned)(uint16_t)px[2 * k] | (unsigned)(uint16_t)px[2 * k + 1] << 16;
    } else {
      d[k] = __builtin_bit_cast(unsigned, px[k]);
    }
  }
}

template <typename TO>
__device__ __forceinline__ void ring8_store_px(TO v, __amdgpu_buffer_rsrc_t rs, int off) {
  if constexpr (sizeof(TO) == 1)
    __builtin_amdgcn_raw_buffer_store_b8((unsigned char)v, rs, off, 0, 2);
  else if constexpr (sizeof(TO) == 2)
    __builtin_amdgcn_raw_buffer_store_b16((unsigned short)v, rs, off, 0, 2);
  else
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), rs, off, 0, 2);
}

// Output rows [y0, y0 + nout) of the strip; xl is the iteration-space column
// of the lane's pixel 0; store_lane: the lane's pixels are stored; used: the
// lane's pixels feed some stored pixel (idle and far lanes load nothing on
// the edge path).
template <typename TO, bool ISUM, int K, bool FMA, int CH, int NS, int ND, bool EDGE>
__device__ __forceinline__ void ring8_body(const LocalArgs& a, uint4* __restrict__ lds,
                                           const uint8_t* __restrict__ in, TO* __restrict__ out,
                                           int xl, bool store_lane, bool used, int y0, int nout) {
  constexpr int V = kRing8V, QR = CH / 4;
  constexpr int R = K / 2;
  constexpr int NT = QR + K - 1;           // LDS rows a wave reads per chunk
  constexpr int BO = (int)sizeof(TO);
  constexpr int NW = V * BO / 4;           // output dwords per lane and row
  static_assert(R <= 4 && 2 * R <= CH, "mask radius exceeds one neighbour word");
  // float weights, or the integer weights as exact floats
  float wt[K * K];
#pragma unroll
  for (int t = 0; t < K * K; ++t) wt[t] = ISUM ? (float)a.w.i[t] : a.w.f[t];
  constexpr int NSLOT = EDGE ? 3 : ND + 2;
  const Geom& g = a.g;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane((threadIdx.x >> 6) & 3);
  const int nin = nout + K - 1;
  const int iy0 = g.in_oy + y0 - R;
  const int ix = g.in_ox + xl;
  const bool cmode = g.boundary == HIPACC_BOUNDARY_CONSTANT;
  const unsigned cb = (unsigned)const_pixel<uint8_t>(g) * 0x01010101u;
  const float rnd = g.cvt == HIPACC_CVT_ROUND ? 0.5f : 0.0f;
  // edge strips: byte offsets of the 16-byte load (raw lanes) and of the
  // per-pixel loads (boundary-mapped columns); cm marks the bytes that take
  // the CONSTANT value
  int roff = kRingDrop;
  int eoff[V];
  unsigned cm[4] = {0u, 0u, 0u, 0u};
  if constexpr (EDGE) {
    int lx, hx, ly, hy;
    raw_region(g, lx, hx, ly, hy);
    const bool lane_raw = a.in_vec_ok && ix >= lx && ix + V <= hx;
    const LaneCols<V> cols = lane_cols<V>(g, ix);
    roff = lane_raw && used ? ix : kRingDrop;
#pragma unroll
    for (int e = 0; e < V; ++e) {
      cm[e / 4] |= (!lane_raw && cols.cst[e]) ? 0xffu << (8 * (e % 4)) : 0u;
      eoff[e] = (lane_raw || cols.cst[e] || !used) ? kRingDrop : cols.x[e];
    }
  }
  const int in_row_bytes = g.in_imgw;

  auto load_chunk = [&](int c) {
    Ring8Stage<QR> st;
    st.rc = 0;
#pragma unroll
    for (int i = 0; i < QR; ++i) {
      const int k = min(c * CH + wave + 4 * i, nin - 1);
      bool rc;
      const int y = row_src(g, iy0 + k, rc);
      st.rc |= rc ? 1 << i : 0;
      const __amdgpu_buffer_rsrc_t rs = ring_rsrc(in + (long long)y * g.in_stride, in_row_bytes);
      st.r[i] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(rs, roff, 0, 0));
#pragma unroll
      for (int e = 0; e < V; ++e)
        st.p[i][e] = (unsigned)__builtin_amdgcn_raw_buffer_load_b8(rs, eoff[e], 0, 0);
    }
    return st;
  };
  auto write_chunk = [&](int c, const Ring8Stage<QR>& st) {
    const int slot = c % NSLOT;
#pragma unroll
    for (int i = 0; i < QR; ++i) {
      unsigned w[4] = {st.r[i].x, st.r[i].y, st.r[i].z, st.r[i].w};
#pragma unroll
      for (int e = 0; e < V; ++e) w[e / 4] |= st.p[i][e] << (8 * (e % 4));
      if (cmode) {
        const bool rc = (st.rc >> i) & 1;
#pragma unroll
        for (int k = 0; k < 4; ++k) w[k] = rc ? cb : ((w[k] & ~cm[k]) | (cb & cm[k]));
      }
      lds[(slot * CH + wave + 4 * i) * 64 + lane] = make_uint4(w[0], w[1], w[2], w[3]);
    }
  };

  const bool full = store_lane && (!EDGE || (a.out_vec_ok && xl + V <= g.is_w));
  const bool part = store_lane && !full;
  const int voff = full ? xl * BO : kRingDrop;
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
      const int slot = (s + c) % NSLOT;
      const uint4 q = lds[(slot * CH + off - c * CH) * 64 + lane];
      const unsigned qw[4] = {q.x, q.y, q.z, q.w};
      const unsigned lw = from_prev_lane(q.w), rw = from_next_lane(q.x);
      float win[V + 2 * R];
#pragma unroll
      for (int k = 0; k < R; ++k) {
        win[k] = ring8_px(lw, 4 - R + k);
        win[R + V + k] = ring8_px(rw, k);
      }
#pragma unroll
      for (int e = 0; e < V; ++e) win[R + e] = ring8_px(qw[e / 4], e % 4);
#pragma unroll
      for (int qq = 0; qq < QR; ++qq) {
        const int dyi = t - qq;
        if (dyi < 0 || dyi >= K) continue;
#pragma unroll
        for (int dx = 0; dx < K; ++dx) {
          const float w = wt[dyi * K + dx];
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
      TO px[V];
      unsigned d[NW];
      ring8_pack<TO, ISUM>(acc[qq], rnd, px, d);
#pragma unroll
      for (int k = 0; k < NW / 4; ++k) {
        const ring_u4 v = {d[4 * k], d[4 * k + 1], d[4 * k + 2], d[4 * k + 3]};
        __builtin_amdgcn_raw_buffer_store_b128(v, rs, voff + 16 * k, 0, 2);   // 2: nontemporal
      }
      if constexpr (EDGE) {
#pragma unroll
        for (int e = 0; e < V; ++e)
          ring8_store_px<TO>(px[e], rs, (part && xl + e < g.is_w) ? (xl + e) * BO : kRingDrop);
      }
    }
  };

  const int nos = (nout + CH - 1) / CH;
  if constexpr (!EDGE) {
    // Direct loads and waits as in ring_body: chunk c lands in slot c % NSLOT.
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
    ring_wait_vm<0>();
  } else {
    Ring8Stage<QR> sets[NS];
    {
      const Ring8Stage<QR> c0 = load_chunk(0), c1 = load_chunk(1);
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

// One item per block of 4 waves (ring_item of hh_ring.hpp); rg.sw stored
// columns per strip.
template <typename TO, bool ISUM, int K, bool FMA, int CH, int NS, int ND>
__global__ __launch_bounds__(256, 3) void ring8_kernel(const LocalArgs a, const RingGeom rg) {
  __shared__ __align__(16) uint4 lds[(ND + 2 > 3 ? ND + 2 : 3) * CH * 64];
  const Geom& g = a.g;
  if ((int)blockIdx.x >= rg.items) return;
  const RingItem it = ring_item(rg, g, blockIdx.x);
  const uint8_t* in = static_cast<const uint8_t*>(g.in) + (long long)it.b * g.in_plane;
  TO* out = static_cast<TO*>(g.out) + (long long)it.b * g.out_plane;
  constexpr int V = kRing8V, R = K / 2;
  const int lane = threadIdx.x & 63;
  const int sl = rg.sw / V;                      // stored lanes
  const int xs = it.strip * rg.sw - V;           // iteration-space column of lane 0's pixel 0
  const int ll = lane <= sl + 1 ? lane : sl + 1; // idle lanes repeat the right halo lane
  const int xl = xs + ll * V;
  const bool store_lane = lane >= 1 && lane <= sl;
  const int s1 = min(xs + V + rg.sw, g.is_w);    // end of the stored columns
  const bool used = lane <= sl + 1 && xl + V > xs + V - R && xl < s1 + R;
  int lx, hx, ly, hy;
  raw_region(g, lx, hx, ly, hy);
  const int ix0 = g.in_ox + xs;
  const int iy0 = g.in_oy + it.y0 - R;
  const bool raw = a.in_vec_ok && a.out_vec_ok && ix0 >= lx && ix0 + (sl + 2) * V <= hx &&
                   xs + (sl + 1) * V <= g.is_w &&
                   (g.boundary != HIPACC_BOUNDARY_CONSTANT ||
                    (iy0 >= g.lo_y && iy0 + it.nout + K - 1 <= g.hi_y));
  if (raw)
    ring8_body<TO, ISUM, K, FMA, CH, NS, ND, false>(a, lds, in, out, xl, store_lane, used, it.y0,
                                                  it.nout);
  else
    ring8_body<TO, ISUM, K, FMA, CH, NS, ND, true>(a, lds, in, out, xl, store_lane, used, it.y0,
                                                 it.nout);
}

}  // namespace hh
