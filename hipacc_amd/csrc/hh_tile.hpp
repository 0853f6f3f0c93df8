// hh_tile.hpp -- LDS staging of an input tile plus halo, and vector I/O.
//
// A block of 256 threads (4 waves) owns a tile of 64*VEC columns by 4*PPT rows
// of the iteration space; VEC = 16 / sizeof(pixel), so every lane moves one
// 16-byte vector per row.  The tile plus HV vectors of halo on each side and
// RY rows above and below is staged in LDS:
//  - interior tiles (the whole staged region lies where raw reads are valid)
//    are loaded with aligned 16-byte global loads and no index arithmetic;
//  - border tiles resolve the boundary mode per staged element.
// This is the "boundary handling only in border blocks" split of the
// reference (runtime/hipacc_cu_standalone.hpp:66-103, lib/AST/
// ASTTranslate.cpp:787-980), applied once while staging instead of per tap.
#pragma once

#include <type_traits>

#include "hh_local.hpp"

namespace hh {

template <int BYTES> struct word_of;
template <> struct word_of<1> { using type = uint8_t; };
template <> struct word_of<2> { using type = uint16_t; };
template <> struct word_of<4> { using type = uint32_t; };
template <> struct word_of<8> { using type = uint2; };
template <> struct word_of<16> { using type = uint4; };

// Stores VEC values of TO at dst, which is aligned to min(16, VEC*sizeof(TO)),
// with nontemporal stores (operator outputs are not re-read by the kernel).
template <typename TO, int VEC>
__device__ __forceinline__ void store_vec(TO* dst, const TO (&v)[VEC]) {
  constexpr int BYTES = VEC * (int)sizeof(TO);
  if constexpr (BYTES >= 16) {
    typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
#pragma unroll
    for (int k = 0; k < BYTES / 16; ++k) {
      u32x4 w;
      __builtin_memcpy(&w, reinterpret_cast<const char*>(v) + 16 * k, 16);
      __builtin_nontemporal_store(w, reinterpret_cast<u32x4*>(dst) + k);
    }
  } else {
    using W = typename std::conditional<
        BYTES == 8, unsigned long long,
        typename std::conditional<BYTES == 4, unsigned int,
                                  typename std::conditional<BYTES == 2, unsigned short,
                                                            unsigned char>::type>::type>::type;
    W w;
    __builtin_memcpy(&w, v, BYTES);
    __builtin_nontemporal_store(w, reinterpret_cast<W*>(dst));
  }
}

template <typename TI> __device__ __forceinline__ TI const_pixel(const Geom& g);
template <> __device__ __forceinline__ float const_pixel<float>(const Geom& g) { return g.const_f; }
template <> __device__ __forceinline__ uint8_t const_pixel<uint8_t>(const Geom& g) { return (uint8_t)g.const_i; }
template <> __device__ __forceinline__ int8_t const_pixel<int8_t>(const Geom& g) { return (int8_t)g.const_i; }
template <> __device__ __forceinline__ uint16_t const_pixel<uint16_t>(const Geom& g) { return (uint16_t)g.const_i; }
template <> __device__ __forceinline__ int16_t const_pixel<int16_t>(const Geom& g) { return (int16_t)g.const_i; }
template <> __device__ __forceinline__ uint32_t const_pixel<uint32_t>(const Geom& g) { return (uint32_t)g.const_i; }
template <> __device__ __forceinline__ int32_t const_pixel<int32_t>(const Geom& g) { return (int32_t)g.const_i; }

// Region where raw (unmapped) reads return what the boundary mode would: the
// accessor window, or the whole allocation for UNDEFINED.
__device__ __forceinline__ void raw_region(const Geom& g, int& lx, int& hx, int& ly, int& hy) {
  if (g.boundary == HIPACC_BOUNDARY_UNDEFINED) {
    lx = 0; hx = g.in_imgw; ly = 0; hy = g.in_imgh;
  } else {
    lx = g.lo_x; hx = g.hi_x; ly = g.lo_y; hy = g.hi_y;
  }
}

// Stages rows [iy0, iy0+LH) x columns [ix0, ix0+NV*VEC) of the input into
// tile (row pitch NV*VEC elements).
template <typename TI, int VEC, int NV, int LH>
__device__ __forceinline__ void stage_tile(TI* __restrict__ tile, const TI* __restrict__ in,
                                           const Geom& g, int ix0, int iy0, int vec_ok) {
  constexpr int LW = NV * VEC;
  const int tid = threadIdx.x;
  int lx, hx, ly, hy;
  raw_region(g, lx, hx, ly, hy);
  const bool interior = vec_ok && ix0 >= lx && ix0 + LW <= hx && iy0 >= ly && iy0 + LH <= hy;
  if (interior) {
    // Every load of the thread is issued before the first LDS write, so all
    // of them are in flight together.
    constexpr int NL = (LH * NV + 255) / 256;
    const TI* src = in + (long long)iy0 * g.in_stride + ix0;
    uint4 w[NL];
#pragma unroll
    for (int k = 0; k < NL; ++k) {
      const int i = tid + k * 256;
      if (i < LH * NV) {
        const int r = i / NV;
        const int c = i - r * NV;
        w[k] = *reinterpret_cast<const uint4*>(src + (long long)r * g.in_stride + c * VEC);
      }
    }
#pragma unroll
    for (int k = 0; k < NL; ++k) {
      const int i = tid + k * 256;
      if (i < LH * NV) {
        const int r = i / NV;
        const int c = i - r * NV;
        *reinterpret_cast<uint4*>(tile + r * LW + c * VEC) = w[k];
      }
    }
  } else {
    const bool cmode = g.boundary == HIPACC_BOUNDARY_CONSTANT;
    const TI cval = const_pixel<TI>(g);
    for (int i = tid; i < LH * LW; i += 256) {
      const int r = i / LW;
      const int c = i - r * LW;
      int x = ix0 + c;
      int y = iy0 + r;
      TI v;
      if (cmode) {
        if (x < g.lo_x || x >= g.hi_x || y < g.lo_y || y >= g.hi_y)
          v = cval;
        else
          v = in[(long long)y * g.in_stride + x];
      } else {
        x = bh_map(g.boundary, x, g.lo_x, g.hi_x, g.in_imgw);
        y = bh_map(g.boundary, y, g.lo_y, g.hi_y, g.in_imgh);
        v = in[(long long)y * g.in_stride + x];
      }
      tile[r * LW + c] = v;
    }
  }
}

// Loads NW consecutive 16-byte vectors of LDS and widens them to ACC.
template <typename TI, typename ACC, int VEC, int NW>
__device__ __forceinline__ void load_window(ACC (&win)[NW * VEC], const TI* __restrict__ p) {
#pragma unroll
  for (int k = 0; k < NW; ++k) {
    TI t[VEC];
    const uint4 w = *reinterpret_cast<const uint4*>(p + k * VEC);
    __builtin_memcpy(t, &w, 16);
#pragma unroll
    for (int v = 0; v < VEC; ++v) win[k * VEC + v] = (ACC)t[v];
  }
}

// Writes one output row segment of VEC pixels (iteration-space-relative
// column ox, row oy) of a tile.
template <typename TO, typename ACC, int VEC>
__device__ __forceinline__ void store_row(TO* __restrict__ out, const Geom& g, int out_vec_ok,
                                          int ox, int oy, const ACC (&acc)[VEC]) {
  if (oy >= g.is_h || ox >= g.is_w) return;
  TO* dst = out + (long long)(g.out_oy + oy) * g.out_stride + g.out_ox + ox;
  TO vals[VEC];
#pragma unroll
  for (int v = 0; v < VEC; ++v) {
    if constexpr (std::is_same<ACC, float>::value)
      vals[v] = convert_f<TO>((float)acc[v], g.cvt, g.post_div);
    else
      vals[v] = convert_i<TO>((int32_t)acc[v], g.cvt, g.post_div);
  }
  if (out_vec_ok && ox + VEC <= g.is_w) {
    store_vec<TO, VEC>(dst, vals);
  } else {
#pragma unroll
    for (int v = 0; v < VEC; ++v)
      if (ox + v < g.is_w) dst[v] = vals[v];
  }
}

}  // namespace hh
