// fused.hip -- fused operator chains (hipacc_hip_fused): point and local
// stages of one iteration-space size in one launch, without the
// intermediate images of the unfused chain and with the same bits.
//
// 64 x 16 output tiles, 256 threads, a column of 4 outputs per lane.
//  no local    each lane runs the point stages on its pixels.
//  one local   the block stages its tile of the local stage's input image
//              (the pre stages applied once per staged pixel, with that
//              image's boundary handling) plus the mask halo in LDS, then
//              evaluates the mask from LDS and runs the post stages.
//  two locals  the block stages the first local's input over the tile plus
//              both halos, evaluates the first local (and the stages between
//              the locals) over the tile plus the second halo into a second
//              LDS tile, then the second local from it.  Where the second
//              local's boundary handling maps a halo pixel to a pixel whose
//              footprint the staged region does not hold (REPEAT, and the
//              halo beyond the right or bottom edge of a partial tile), that
//              pixel of the intermediate image is evaluated from global
//              memory instead, with the same arithmetic.
// Every value is produced by the arithmetic of the unfused operator:
// point_pixel (hipacc_hip_point), the generic local operator's tap order,
// seeding and conversions, and each intermediate image's pixel type.
#include "hh_fused.hpp"
#include "hh_point.hpp"

namespace hh {

constexpr int kFusedTW = 64, kFusedTH = 16;

// A pixel of an image of pixel type dt: exact as float (F32) or int32.
struct FVal {
  float f;
  int32_t i;
};

__device__ __forceinline__ FVal fv_load(const void* p, int dt, long long idx) {
  FVal v{0.0f, 0};
  if (dt == HIPACC_F32) v.f = load_f(p, HIPACC_F32, idx);
  else v.i = load_i(p, dt, idx);
  return v;
}

template <typename TO> __device__ __forceinline__ FVal fv_of(TO r) {
  FVal v{0.0f, 0};
  if constexpr (std::is_same<TO, float>::value) v.f = r;
  else v.i = (int32_t)r;
  return v;
}

template <typename TO, typename ACC>
__device__ __forceinline__ TO convert_acc(ACC acc, int cvt, float post_div) {
  if constexpr (std::is_same<ACC, float>::value) return convert_f<TO>(acc, cvt, post_div);
  else return convert_i<TO>(acc, cvt, post_div);
}

// A float (or int) result stored to pixel type dt, as the generic local
// operator's store_f / store_i.
template <typename ACC>
__device__ __forceinline__ FVal fv_convert(int dt, ACC acc, int cvt, float post_div) {
  switch (dt) {
    case HIPACC_U8: return fv_of(convert_acc<uint8_t>(acc, cvt, post_div));
    case HIPACC_I8: return fv_of(convert_acc<int8_t>(acc, cvt, post_div));
    case HIPACC_U16: return fv_of(convert_acc<uint16_t>(acc, cvt, post_div));
    case HIPACC_I16: return fv_of(convert_acc<int16_t>(acc, cvt, post_div));
    case HIPACC_I32: return fv_of(convert_acc<int32_t>(acc, cvt, post_div));
    default: return fv_of(convert_acc<float>(acc, cvt, post_div));
  }
}

template <bool FM>
__device__ __forceinline__ FVal point_to(const FusedPoint& s, float af, float bf, int32_t ai,
                                         int32_t bi) {
  switch (s.dtype) {
    case HIPACC_U8: return fv_of(point_pixel<uint8_t, FM>(s.op, af, bf, ai, bi, s.alpha, s.beta, s.cvt));
    case HIPACC_I8: return fv_of(point_pixel<int8_t, FM>(s.op, af, bf, ai, bi, s.alpha, s.beta, s.cvt));
    case HIPACC_U16: return fv_of(point_pixel<uint16_t, FM>(s.op, af, bf, ai, bi, s.alpha, s.beta, s.cvt));
    case HIPACC_I16: return fv_of(point_pixel<int16_t, FM>(s.op, af, bf, ai, bi, s.alpha, s.beta, s.cvt));
    case HIPACC_I32: return fv_of(point_pixel<int32_t, FM>(s.op, af, bf, ai, bi, s.alpha, s.beta, s.cvt));
    default: return fv_of(point_pixel<float, FM>(s.op, af, bf, ai, bi, s.alpha, s.beta, s.cvt));
  }
}

// Point stages [from, to) on value v of pixel type dt at pixel (x, y) of
// image b; dt becomes the last stage's pixel type.
__device__ __forceinline__ FVal run_points(const FusedArgs& a, int from, int to, int& dt, FVal v,
                                           int x, int y, int b) {
  for (int k = from; k < to; ++k) {
    const FusedPoint& s = a.pt[k];
    FVal u{0.0f, 0};
    if (s.in1.data) {
      const char* p = static_cast<const char*>(s.in1.data) +
                      (long long)b * s.in1.plane * dtype_bytes(s.in1.dtype);
      u = fv_load(p, s.in1.dtype, (long long)(s.in1.oy + y) * s.in1.stride + s.in1.ox + x);
    }
    if (dt == HIPACC_F32)
      v = point_to<true>(s, v.f, u.f, (int32_t)v.f, (int32_t)u.f);
    else
      v = point_to<false>(s, (float)v.i, (float)u.i, v.i, u.i);
    dt = s.dtype;
  }
  return v;
}

template <typename ACC> __device__ __forceinline__ ACC to_acc(FVal v, int dt) {
  if constexpr (std::is_same<ACC, float>::value) return dt == HIPACC_F32 ? v.f : (float)v.i;
  else return dt == HIPACC_F32 ? (int32_t)v.f : v.i;
}

__device__ __forceinline__ void fv_store(void* out, int dt, long long o, FVal v) {
  switch (dt) {
    case HIPACC_U8: static_cast<uint8_t*>(out)[o] = (uint8_t)v.i; break;
    case HIPACC_I8: static_cast<int8_t*>(out)[o] = (int8_t)v.i; break;
    case HIPACC_U16: static_cast<uint16_t*>(out)[o] = (uint16_t)v.i; break;
    case HIPACC_I16: static_cast<int16_t*>(out)[o] = (int16_t)v.i; break;
    case HIPACC_I32: static_cast<int32_t*>(out)[o] = v.i; break;
    default: static_cast<float*>(out)[o] = v.f; break;
  }
}

// The mask of local L evaluated with the generic local operator's
// arithmetic; fetch(dx, dy) yields the input tap at mask offset (dx, dy).
template <typename ACC, typename F>
__device__ __forceinline__ ACC local_eval(const FusedLocal& L, F&& fetch) {
  constexpr bool IS_F = std::is_same<ACC, float>::value;
  ACC acc = ACC(0);
  bool first = true;
  for (int dy = 0; dy < L.mask_h; ++dy)
    for (int dx = 0; dx < L.mask_w; ++dx) {
      const int t = dy * L.mask_w + dx;
      if (!L.domain[t]) continue;
      const ACC v = fetch(dx, dy);
      ACC w = ACC(1);
      if (L.use_weights) {
        if constexpr (IS_F) w = L.w.f[t];
        else w = (ACC)L.w.i[t];
      }
      if (L.reduce == HIPACC_REDUCE_SUM && L.use_weights) {
        if (first) acc = madd<false>(w, v, ACC(0));
        else if (L.fma) acc = madd<true>(w, v, acc);
        else acc = madd<false>(w, v, acc);
      } else {
        ACC tap = L.use_weights ? (ACC)(w * v) : v;
        if constexpr (!IS_F)
          if (L.use_weights) tap = (ACC)((uint32_t)w * (uint32_t)v);
        acc = first ? tap : reduce_combine<ACC>(L.reduce, acc, tap);
      }
      first = false;
    }
  return acc;
}

__device__ __forceinline__ bool outside(const FusedArgs& a, int x, int y) {
  return x < 0 || x >= a.w || y < 0 || y >= a.h;
}

// Pixel (x, y) (iteration coordinates, possibly outside) of the first
// local's input image as seen through its boundary handling: `in` itself
// when no point stage precedes the local (its window and allocation, as
// hipacc_hip_local), else the image the pre stages write.
template <typename ACC>
__device__ __forceinline__ ACC first_input(const FusedArgs& a, const char* in, int x, int y, int b) {
  const FusedLocal& L = a.loc[0];
  if (L.boundary == HIPACC_BOUNDARY_CONSTANT) {
    if (outside(a, x, y)) return std::is_same<ACC, float>::value ? (ACC)L.const_f : (ACC)L.const_i;
  } else if (a.seg[0][1] == a.seg[0][0]) {
    x = bh_map(L.boundary, x + a.in_ox, a.in_ox, a.in_ox + a.w, a.in_imgw) - a.in_ox;
    y = bh_map(L.boundary, y + a.in_oy, a.in_oy, a.in_oy + a.h, a.in_imgh) - a.in_oy;
  } else {
    x = bh_map(L.boundary, x, 0, a.w, a.w);
    y = bh_map(L.boundary, y, 0, a.h, a.h);
  }
  int dt = a.in_dtype;
  FVal v = fv_load(in, dt, (long long)(a.in_oy + y) * a.in_stride + a.in_ox + x);
  v = run_points(a, a.seg[0][0], a.seg[0][1], dt, v, x, y, b);
  return to_acc<ACC>(v, dt);
}

template <typename ACC1, typename ACC2, int NL>
__global__ __launch_bounds__(256) void fused_kernel(const FusedArgs a) {
  const int x0 = blockIdx.x * kFusedTW, y0 = blockIdx.y * kFusedTH;
  const int b = blockIdx.z;
  const char* in = static_cast<const char*>(a.in) +
                   (long long)b * a.in_plane * dtype_bytes(a.in_dtype);
  char* out = static_cast<char*>(a.out) + (long long)b * a.out_plane * dtype_bytes(a.out_dtype);
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int x = x0 + tx;
  const int ps = a.seg[2][0], pe = a.seg[2][1];
  if constexpr (NL == 0) {
    if (x >= a.w) return;
#pragma unroll 1
    for (int j = 0; j < 4; ++j) {
      const int y = y0 + ty * 4 + j;
      if (y >= a.h) return;
      int dt = a.in_dtype;
      FVal v = fv_load(in, dt, (long long)(a.in_oy + y) * a.in_stride + a.in_ox + x);
      v = run_points(a, ps, pe, dt, v, x, y, b);
      fv_store(out, dt, (long long)(a.out_oy + y) * a.out_stride + a.out_ox + x, v);
    }
  } else {
    extern __shared__ __align__(16) unsigned char fused_lds[];
    const FusedLocal& L0 = a.loc[0];
    const FusedLocal& LL = a.loc[NL - 1];   // the local the outputs come from
    const int rxl = LL.mask_w / 2, ryl = LL.mask_h / 2;
    const int W2 = kFusedTW + LL.mask_w - 1, H2 = kFusedTH + LL.mask_h - 1;
    // Tile of the last local's input, origin (x0 - rxl, y0 - ryl).
    ACC2* tile2;
    if constexpr (NL == 1) {
      ACC1* t = reinterpret_cast<ACC1*>(fused_lds);
      for (int i = threadIdx.x; i < W2 * H2; i += 256) {
        const int r = i / W2, c = i - r * W2;
        t[i] = first_input<ACC1>(a, in, x0 - rxl + c, y0 - ryl + r, b);
      }
      tile2 = reinterpret_cast<ACC2*>(t);
    } else {
      // Tile of the first local's input over the tile plus both halos,
      // origin (x0 - rxl - rx0, y0 - ryl - ry0).
      const int rx0 = L0.mask_w / 2, ry0 = L0.mask_h / 2;
      const int W1 = W2 + L0.mask_w - 1, H1 = H2 + L0.mask_h - 1;
      ACC1* t1 = reinterpret_cast<ACC1*>(fused_lds);
      tile2 = reinterpret_cast<ACC2*>(fused_lds + (((size_t)W1 * H1 * 4 + 15) & ~(size_t)15));
      for (int i = threadIdx.x; i < W1 * H1; i += 256) {
        const int r = i / W1, c = i - r * W1;
        t1[i] = first_input<ACC1>(a, in, x0 - rxl - rx0 + c, y0 - ryl - ry0 + r, b);
      }
      __syncthreads();
      const int ox2 = x0 - rxl, oy2 = y0 - ryl;
      for (int i = threadIdx.x; i < W2 * H2; i += 256) {
        const int r = i / W2, c = i - r * W2;
        int qx = ox2 + c, qy = oy2 + r;
        ACC2 v2;
        if (LL.boundary == HIPACC_BOUNDARY_CONSTANT && outside(a, qx, qy)) {
          v2 = std::is_same<ACC2, float>::value ? (ACC2)LL.const_f : (ACC2)LL.const_i;
        } else {
          if (LL.boundary != HIPACC_BOUNDARY_CONSTANT) {
            qx = bh_map(LL.boundary, qx, 0, a.w, a.w);
            qy = bh_map(LL.boundary, qy, 0, a.h, a.h);
          }
          const int lc = qx - ox2, lr = qy - oy2;
          ACC1 acc;
          if (lc >= 0 && lc < W2 && lr >= 0 && lr < H2) {
            acc = local_eval<ACC1>(L0, [&](int dx, int dy) { return t1[(lr + dy) * W1 + lc + dx]; });
          } else {
            acc = local_eval<ACC1>(L0, [&](int dx, int dy) {
              return first_input<ACC1>(a, in, qx + dx - rx0, qy + dy - ry0, b);
            });
          }
          int dt = L0.out_dtype;
          FVal v = fv_convert(dt, acc, L0.cvt, L0.post_div);
          v = run_points(a, a.seg[1][0], a.seg[1][1], dt, v, qx, qy, b);
          v2 = to_acc<ACC2>(v, dt);
        }
        tile2[i] = v2;
      }
    }
    __syncthreads();
#pragma unroll 1
    for (int j = 0; j < 4; ++j) {
      const int yl = ty * 4 + j, y = y0 + yl;
      if (x >= a.w || y >= a.h) break;
      const ACC2 acc = local_eval<ACC2>(LL, [&](int dx, int dy) { return tile2[(yl + dy) * W2 + tx + dx]; });
      int dt = LL.out_dtype;
      FVal v = fv_convert(dt, acc, LL.cvt, LL.post_div);
      v = run_points(a, ps, pe, dt, v, x, y, b);
      fv_store(out, dt, (long long)(a.out_oy + y) * a.out_stride + a.out_ox + x, v);
    }
  }
}

// ---------------------------------------------------------------------------
// Fast path: every image of the chain of one pixel type T (float or uint8),
// square K x K locals (3, 5, 7) that are weighted float SUMs over the whole
// mask with one contraction flag.  Blocks of 64 - (KL - 1) output columns
// (KL: the last local's width) by 32 rows, so that the last local's input
// tile is exactly 64 columns: lane c of each wave owns column c of it.  A
// lane evaluates a mask down a column of RN outputs, reading each staged
// row once for all the outputs it feeds (tap order per output unchanged:
// rows, then columns, first tap seeding the sum).
// ---------------------------------------------------------------------------
template <typename T> using FastV = typename std::conditional<std::is_same<T, float>::value,
                                                              float, int32_t>::type;

template <typename T>
__device__ __forceinline__ FastV<T> points_t(const FusedArgs& a, int from, int to, FastV<T> v,
                                             int x, int y, int b) {
  for (int k = from; k < to; ++k) {
    const FusedPoint& s = a.pt[k];
    FastV<T> u = 0;
    if (s.in1.data)
      u = (FastV<T>)static_cast<const T*>(s.in1.data)[(long long)b * s.in1.plane +
                                                      (long long)(s.in1.oy + y) * s.in1.stride +
                                                      s.in1.ox + x];
    if constexpr (std::is_same<T, float>::value)
      v = point_pixel<float, true>(s.op, v, u, (int32_t)v, (int32_t)u, s.alpha, s.beta, s.cvt);
    else
      v = (int32_t)point_pixel<T, false>(s.op, (float)v, (float)u, v, u, s.alpha, s.beta, s.cvt);
  }
  return v;
}

// points_t over N values at once, stage by stage: each stage's parameters
// and operator are decoded once per wave for all N values instead of once
// per value.  xy(e) gives value e's pixel (inside the image: it is where a
// second input is read).
template <typename T, int OP, int N>
__device__ __forceinline__ void point_op_vec(const FusedPoint& s, FastV<T> (&v)[N],
                                             const FastV<T> (&u)[N]) {
#pragma unroll
  for (int e = 0; e < N; ++e) {
    if constexpr (std::is_same<T, float>::value)
      v[e] = point_pixel<float, true>(OP, v[e], u[e], (int32_t)v[e], (int32_t)u[e], s.alpha,
                                      s.beta, s.cvt);
    else
      v[e] = (int32_t)point_pixel<T, false>(OP, (float)v[e], (float)u[e], v[e], u[e], s.alpha,
                                            s.beta, s.cvt);
  }
}

template <typename T, int N, typename XY>
__device__ __forceinline__ void points_vec(const FusedArgs& a, int from, int to,
                                           FastV<T> (&v)[N], XY&& xy, int b) {
  for (int k = from; k < to; ++k) {
    const FusedPoint& s = a.pt[k];
    FastV<T> u[N];
    if (s.in1.data) {
      const T* p = static_cast<const T*>(s.in1.data) + (long long)b * s.in1.plane;
#pragma unroll
      for (int e = 0; e < N; ++e) {
        const int2 q = xy(e);
        u[e] = (FastV<T>)p[(long long)(s.in1.oy + q.y) * s.in1.stride + s.in1.ox + q.x];
      }
    } else {
#pragma unroll
      for (int e = 0; e < N; ++e) u[e] = 0;
    }
    switch (s.op) {
      case HIPACC_POINT_COPY: point_op_vec<T, HIPACC_POINT_COPY, N>(s, v, u); break;
      case HIPACC_POINT_AFFINE: point_op_vec<T, HIPACC_POINT_AFFINE, N>(s, v, u); break;
      case HIPACC_POINT_ADD: point_op_vec<T, HIPACC_POINT_ADD, N>(s, v, u); break;
      case HIPACC_POINT_SUB: point_op_vec<T, HIPACC_POINT_SUB, N>(s, v, u); break;
      case HIPACC_POINT_MUL: point_op_vec<T, HIPACC_POINT_MUL, N>(s, v, u); break;
      case HIPACC_POINT_ABSDIFF: point_op_vec<T, HIPACC_POINT_ABSDIFF, N>(s, v, u); break;
      case HIPACC_POINT_MIN: point_op_vec<T, HIPACC_POINT_MIN, N>(s, v, u); break;
      case HIPACC_POINT_MAX: point_op_vec<T, HIPACC_POINT_MAX, N>(s, v, u); break;
      default: point_op_vec<T, HIPACC_POINT_MAGNITUDE, N>(s, v, u); break;
    }
  }
}

// first_input for the fast path: the same pixel, as float.
template <typename T>
__device__ __forceinline__ float first_input_t(const FusedArgs& a, const T* in, int x, int y, int b) {
  const FusedLocal& L = a.loc[0];
  if (L.boundary == HIPACC_BOUNDARY_CONSTANT) {
    if (outside(a, x, y)) return L.const_f;
  } else if (a.seg[0][1] == a.seg[0][0]) {
    x = bh_map(L.boundary, x + a.in_ox, a.in_ox, a.in_ox + a.w, a.in_imgw) - a.in_ox;
    y = bh_map(L.boundary, y + a.in_oy, a.in_oy, a.in_oy + a.h, a.in_imgh) - a.in_oy;
  } else {
    x = bh_map(L.boundary, x, 0, a.w, a.w);
    y = bh_map(L.boundary, y, 0, a.h, a.h);
  }
  FastV<T> v = (FastV<T>)in[(long long)(a.in_oy + y) * a.in_stride + a.in_ox + x];
  return (float)points_t<T>(a, a.seg[0][0], a.seg[0][1], v, x, y, b);
}

// Stages the W x H tile of the first local's input whose element (0, 0) is
// iteration-space pixel (ox, oy).  A tile inside the image needs no boundary
// handling: every thread issues all its loads before the first LDS write.
template <typename T, int W, int H>
__device__ __forceinline__ void stage_tile(const FusedArgs& a, const T* in, float* tile, int ox,
                                           int oy, int b) {
  constexpr int N = W * H, NPT = (N + 255) / 256;
  if (ox >= 0 && oy >= 0 && ox + W <= a.w && oy + H <= a.h) {
    const T* base = in + (long long)(a.in_oy + oy) * a.in_stride + a.in_ox + ox;
    FastV<T> v[NPT];
#pragma unroll
    for (int k = 0; k < NPT; ++k) {
      const int i = min((int)threadIdx.x + 256 * k, N - 1);
      const int r = i / W, c = i - r * W;
      v[k] = (FastV<T>)base[(long long)r * a.in_stride + c];
    }
    points_vec<T, NPT>(a, a.seg[0][0], a.seg[0][1], v, [&](int k) {
      const int i = min((int)threadIdx.x + 256 * k, N - 1);
      const int r = i / W;
      return make_int2(ox + i - r * W, oy + r);
    }, b);
#pragma unroll
    for (int k = 0; k < NPT; ++k) {
      const int i = threadIdx.x + 256 * k;
      if (i < N) tile[i] = (float)v[k];
    }
  } else {
    for (int i = threadIdx.x; i < N; i += 256) {
      const int r = i / W, c = i - r * W;
      tile[i] = first_input_t<T>(a, in, ox + c, oy + r, b);
    }
  }
}

// RN outputs down column `col` of an LDS tile of row pitch W, output j at
// tile rows row0 + j .. row0 + j + K - 1.
template <int K, int RN, bool FMA>
__device__ __forceinline__ void conv_col(const float* tile, int W, int col, int row0,
                                         const float* w, float (&acc)[RN]) {
#pragma unroll
  for (int r = 0; r < RN + K - 1; ++r) {
    float v[K];
#pragma unroll
    for (int dx = 0; dx < K; ++dx) v[dx] = tile[(row0 + r) * W + col + dx];
#pragma unroll
    for (int j = 0; j < RN; ++j) {
      const int dy = r - j;
      if (dy < 0 || dy >= K) continue;
#pragma unroll
      for (int dx = 0; dx < K; ++dx) {
        const int t = dy * K + dx;
        acc[j] = t == 0 ? madd<false>(w[t], v[dx], 0.0f) : madd<FMA>(w[t], v[dx], acc[j]);
      }
    }
  }
}

// One output of a K x K weighted sum; fetch(dx, dy) yields the tap.
template <int K, bool FMA, typename F>
__device__ __forceinline__ float conv_one(const float* w, F&& fetch) {
  float acc = 0.0f;
#pragma unroll
  for (int dy = 0; dy < K; ++dy)
#pragma unroll
    for (int dx = 0; dx < K; ++dx) {
      const int t = dy * K + dx;
      acc = t == 0 ? madd<false>(w[t], fetch(dx, dy), 0.0f) : madd<FMA>(w[t], fetch(dx, dy), acc);
    }
  return acc;
}

// The first local at (qx, qy) with its taps read through first_input_t
// (pixels whose footprint the staged tile does not hold); not unrolled.
template <typename T, int K, bool FMA>
__device__ __forceinline__ float conv_gather(const FusedArgs& a, const T* in, int qx, int qy, int b) {
  constexpr int R = K / 2;
  float acc = 0.0f;
#pragma unroll 1
  for (int t = 0; t < K * K; ++t) {
    const int dy = t / K, dx = t - dy * K;
    const float v = first_input_t<T>(a, in, qx + dx - R, qy + dy - R, b);
    const float w = a.loc[0].w.f[t];
    acc = t == 0 ? madd<false>(w, v, 0.0f) : madd<FMA>(w, v, acc);
  }
  return acc;
}

constexpr int kFastRPW = 8;                 // output rows per wave
constexpr int kFastTH = 4 * kFastRPW;       // output rows per block
template <int K> constexpr int fast_tw() { return 64 - (K - 1); }

template <int K0, int K1> struct FastShape {
  static constexpr int KL = K1 ? K1 : K0;           // the last local
  static constexpr int TW = fast_tw<KL>();
  static constexpr int W2 = 64, H2 = kFastTH + KL - 1;
  static constexpr int R2 = (H2 + 3) / 4;           // intermediate rows per wave
  static constexpr int W1 = W2 + K0 - 1;
  static constexpr int H1 = 4 * R2 + K0 - 1;        // rows the waves read (padded)
  static constexpr int T1 = K1 ? ((W1 * H1 + 3) & ~3) : 0;   // floats of the first tile
  static constexpr size_t bytes = 4 * ((size_t)T1 + W2 * H2);
};

template <typename T, int K0, int K1, bool FMA>
__global__ __launch_bounds__(256) void fused_fast_kernel(const FusedArgs a) {
  using S = FastShape<K0, K1>;
  constexpr int KL = S::KL, RL = KL / 2, TW = S::TW, W2 = S::W2, H2 = S::H2;
  const int x0 = blockIdx.x * TW, y0 = blockIdx.y * kFastTH, b = blockIdx.z;
  const T* in = static_cast<const T*>(a.in) + (long long)b * a.in_plane;
  T* out = static_cast<T*>(a.out) + (long long)b * a.out_plane;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  extern __shared__ __align__(16) float fast_lds[];
  float* t2 = fast_lds + S::T1;
  const FusedLocal& LL = a.loc[K1 ? 1 : 0];
  if constexpr (K1 == 0) {
    stage_tile<T, W2, H2>(a, in, t2, x0 - RL, y0 - RL, b);
  } else {
    constexpr int R0 = K0 / 2, W1 = S::W1, H1 = S::H1, R2 = S::R2;
    const FusedLocal& L0 = a.loc[0];
    float* t1 = fast_lds;
    stage_tile<T, W1, H1>(a, in, t1, x0 - RL - R0, y0 - RL - R0, b);
    __syncthreads();
    const int ox2 = x0 - RL, oy2 = y0 - RL;
    if (ox2 >= 0 && oy2 >= 0 && ox2 + W2 <= a.w && oy2 + H2 <= a.h) {
      // every pixel of the tile is the intermediate image's own pixel
      float acc[R2];
      conv_col<K0, R2, FMA>(t1, W1, lane, wave * R2, L0.w.f, acc);
      FastV<T> v[R2];
#pragma unroll
      for (int j = 0; j < R2; ++j) v[j] = (FastV<T>)convert_f<T>(acc[j], L0.cvt, L0.post_div);
      points_vec<T, R2>(a, a.seg[1][0], a.seg[1][1], v, [&](int j) {
        return make_int2(ox2 + lane, oy2 + min(wave * R2 + j, H2 - 1));
      }, b);
#pragma unroll
      for (int j = 0; j < R2; ++j) {
        const int r = wave * R2 + j;
        if (r < H2) t2[r * W2 + lane] = (float)v[j];
      }
    } else {
      for (int i = threadIdx.x; i < W2 * H2; i += 256) {
        const int r = i >> 6, c = i & 63;
        int qx = ox2 + c, qy = oy2 + r;
        float v2;
        if (LL.boundary == HIPACC_BOUNDARY_CONSTANT && outside(a, qx, qy)) {
          v2 = LL.const_f;
        } else {
          if (LL.boundary != HIPACC_BOUNDARY_CONSTANT) {
            qx = bh_map(LL.boundary, qx, 0, a.w, a.w);
            qy = bh_map(LL.boundary, qy, 0, a.h, a.h);
          }
          const int lc = qx - ox2, lr = qy - oy2;
          float acc;
          if (lc >= 0 && lc < W2 && lr >= 0 && lr < H2)
            acc = conv_one<K0, FMA>(L0.w.f, [&](int dx, int dy) { return t1[(lr + dy) * W1 + lc + dx]; });
          else
            acc = conv_gather<T, K0, FMA>(a, in, qx, qy, b);
          const FastV<T> v = (FastV<T>)convert_f<T>(acc, L0.cvt, L0.post_div);
          v2 = (float)points_t<T>(a, a.seg[1][0], a.seg[1][1], v, qx, qy, b);
        }
        t2[i] = v2;
      }
    }
  }
  __syncthreads();
  float acc[kFastRPW];
  // lanes past the stored columns compute column TW - 1 again and store nothing
  conv_col<KL, kFastRPW, FMA>(t2, W2, lane < TW ? lane : TW - 1, wave * kFastRPW, LL.w.f, acc);
  const int x = x0 + lane;
  if (lane >= TW || x >= a.w) return;
  FastV<T> v[kFastRPW];
#pragma unroll
  for (int j = 0; j < kFastRPW; ++j) v[j] = (FastV<T>)convert_f<T>(acc[j], LL.cvt, LL.post_div);
  points_vec<T, kFastRPW>(a, a.seg[2][0], a.seg[2][1], v, [&](int j) {
    return make_int2(x, min(y0 + wave * kFastRPW + j, a.h - 1));
  }, b);
#pragma unroll
  for (int j = 0; j < kFastRPW; ++j) {
    const int y = y0 + wave * kFastRPW + j;
    if (y < a.h) out[(long long)(a.out_oy + y) * a.out_stride + a.out_ox + x] = (T)v[j];
  }
}

// Whether the chain takes the fast path; its pixel type and mask sizes.
static bool fast_chain(const FusedArgs& a, int& dt, int& k0, int& k1) {
  dt = a.in_dtype;
  if ((dt != HIPACC_F32 && dt != HIPACC_U8) || a.out_dtype != dt || a.num_locals < 1) return false;
  for (int s = 0; s < 3; ++s)
    for (int k = a.seg[s][0]; k < a.seg[s][1]; ++k)
      if (a.pt[k].dtype != dt) return false;
  for (int l = 0; l < a.num_locals; ++l) {
    const FusedLocal& L = a.loc[l];
    if (L.mask_w != L.mask_h || (L.mask_w != 3 && L.mask_w != 5 && L.mask_w != 7) || !L.acc_f ||
        L.reduce != HIPACC_REDUCE_SUM || !L.use_weights || L.in_dtype != dt || L.out_dtype != dt ||
        L.fma != a.loc[0].fma)
      return false;
    for (int t = 0; t < L.mask_w * L.mask_h; ++t)
      if (!L.domain[t]) return false;
  }
  k0 = a.loc[0].mask_w;
  k1 = a.num_locals == 2 ? a.loc[1].mask_w : 0;
  return true;
}

template <typename T, int K0, int K1>
static int launch_fast_k(const FusedArgs& a, hipStream_t s) {
  using S = FastShape<K0, K1>;
  const dim3 grid(ceil_div(a.w, S::TW), ceil_div(a.h, kFastTH), a.batch);
  if (a.loc[0].fma)
    hipLaunchKernelGGL((fused_fast_kernel<T, K0, K1, true>), grid, dim3(256), S::bytes, s, a);
  else
    hipLaunchKernelGGL((fused_fast_kernel<T, K0, K1, false>), grid, dim3(256), S::bytes, s, a);
  return hip_status(hipGetLastError(), "fused_fast_kernel launch");
}
template <typename T, int K0>
static int launch_fast_k1(const FusedArgs& a, int k1, hipStream_t s) {
  switch (k1) {
    case 0: return launch_fast_k<T, K0, 0>(a, s);
    case 3: return launch_fast_k<T, K0, 3>(a, s);
    case 5: return launch_fast_k<T, K0, 5>(a, s);
    default: return launch_fast_k<T, K0, 7>(a, s);
  }
}
template <typename T> static int launch_fast_t(const FusedArgs& a, int k0, int k1, hipStream_t s) {
  switch (k0) {
    case 3: return launch_fast_k1<T, 3>(a, k1, s);
    case 5: return launch_fast_k1<T, 5>(a, k1, s);
    default: return launch_fast_k1<T, 7>(a, k1, s);
  }
}

static size_t fused_lds_bytes(const FusedArgs& a) {
  if (a.num_locals == 0) return 0;
  const FusedLocal& LL = a.loc[a.num_locals - 1];
  const size_t W2 = kFusedTW + LL.mask_w - 1, H2 = kFusedTH + LL.mask_h - 1;
  if (a.num_locals == 1) return 4 * W2 * H2;
  const size_t W1 = W2 + a.loc[0].mask_w - 1, H1 = H2 + a.loc[0].mask_h - 1;
  return ((4 * W1 * H1 + 15) & ~(size_t)15) + 4 * W2 * H2;
}

template <int NL> static int launch_nl(const FusedArgs& a, dim3 grid, size_t smem, hipStream_t s) {
  const bool f0 = a.loc[0].acc_f, f1 = a.loc[NL > 1 ? 1 : 0].acc_f;
  if (NL == 1) {
    if (f0) hipLaunchKernelGGL((fused_kernel<float, float, NL>), grid, dim3(256), smem, s, a);
    else hipLaunchKernelGGL((fused_kernel<int32_t, int32_t, NL>), grid, dim3(256), smem, s, a);
  } else if (f0 && f1) {
    hipLaunchKernelGGL((fused_kernel<float, float, NL>), grid, dim3(256), smem, s, a);
  } else if (f0) {
    hipLaunchKernelGGL((fused_kernel<float, int32_t, NL>), grid, dim3(256), smem, s, a);
  } else if (f1) {
    hipLaunchKernelGGL((fused_kernel<int32_t, float, NL>), grid, dim3(256), smem, s, a);
  } else {
    hipLaunchKernelGGL((fused_kernel<int32_t, int32_t, NL>), grid, dim3(256), smem, s, a);
  }
  return hip_status(hipGetLastError(), "fused_kernel launch");
}

int launch_fused(const FusedArgs& a, hipStream_t s) {
  int dt, k0, k1;
  if (fast_chain(a, dt, k0, k1))
    return dt == HIPACC_F32 ? launch_fast_t<float>(a, k0, k1, s) : launch_fast_t<uint8_t>(a, k0, k1, s);
  const dim3 grid(ceil_div(a.w, kFusedTW), ceil_div(a.h, kFusedTH), a.batch);
  const size_t smem = fused_lds_bytes(a);
  switch (a.num_locals) {
    case 0:
      hipLaunchKernelGGL((fused_kernel<float, float, 0>), grid, dim3(256), 0, s, a);
      return hip_status(hipGetLastError(), "fused_kernel launch");
    case 1: return launch_nl<1>(a, grid, smem, s);
    default: return launch_nl<2>(a, grid, smem, s);
  }
}

}  // namespace hh
