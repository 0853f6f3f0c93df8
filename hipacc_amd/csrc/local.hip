// local.hip -- generic local operator (any mask size up to 15x15, any
// reduction mode, any pixel types, domain holes honoured exactly) and the
// dispatch between it and the templated fast SUM kernels.
#include "hh_tile.hpp"

namespace hh {

// One output pixel per thread; blocks of 64x4.  Blocks whose whole footprint
// reads raw pixels skip boundary handling (uniform branch).
template <typename ACC>
__global__ __launch_bounds__(256) void local_generic_kernel(const LocalArgs a) {
  const Geom& g = a.g;
  const int rx = a.mask_w / 2, ry = a.mask_h / 2;
  const int bx = blockIdx.x * 64, by = blockIdx.y * 4;
  const int x = bx + threadIdx.x, y = by + threadIdx.y;
  int lx, hx, ly, hy;
  raw_region(g, lx, hx, ly, hy);
  const bool interior = g.in_ox + bx - rx >= lx && g.in_ox + bx + 64 + rx <= hx &&
                        g.in_oy + by - ry >= ly && g.in_oy + by + 4 + ry <= hy;
  if (x >= g.is_w || y >= g.is_h) return;
  const char* in = static_cast<const char*>(g.in) +
                   (long long)blockIdx.z * g.in_plane * dtype_bytes(g.in_dtype);
  char* out = static_cast<char*>(g.out) +
              (long long)blockIdx.z * g.out_plane * dtype_bytes(g.out_dtype);
  const int ix = g.in_ox + x, iy = g.in_oy + y;
  const bool cmode = g.boundary == HIPACC_BOUNDARY_CONSTANT;
  const bool is_float = std::is_same<ACC, float>::value;
  ACC acc = ACC(0);
  bool first = true;
  for (int dy = 0; dy < a.mask_h; ++dy) {
    int yy = iy + dy - ry;
    bool row_out = false;
    if (!interior) {
      if (cmode)
        row_out = yy < g.lo_y || yy >= g.hi_y;
      else
        yy = bh_map(g.boundary, yy, g.lo_y, g.hi_y, g.in_imgh);
    }
    for (int dx = 0; dx < a.mask_w; ++dx) {
      const int t = dy * a.mask_w + dx;
      if (!a.domain[t]) continue;
      int xx = ix + dx - rx;
      ACC v;
      bool use_const = row_out;
      if (!interior) {
        if (cmode)
          use_const = use_const || xx < g.lo_x || xx >= g.hi_x;
        else
          xx = bh_map(g.boundary, xx, g.lo_x, g.hi_x, g.in_imgw);
      }
      if (use_const) {
        v = is_float ? (ACC)g.const_f : (ACC)g.const_i;
      } else {
        const long long idx = (long long)yy * g.in_stride + xx;
        if constexpr (std::is_same<ACC, float>::value)
          v = load_f(in, g.in_dtype, idx);
        else
          v = load_i(in, g.in_dtype, idx);
      }
      ACC w = ACC(1);
      if (a.use_weights) {
        if (a.wdev)
          w = (ACC)a.wdev[t];
        else if constexpr (std::is_same<ACC, float>::value)
          w = a.w.f[t];
        else
          w = (ACC)a.w.i[t];
      }
      if (a.reduce == HIPACC_REDUCE_SUM && a.use_weights) {
        if (first)
          acc = madd<false>(w, v, ACC(0));
        else if (a.fma)
          acc = madd<true>(w, v, acc);
        else
          acc = madd<false>(w, v, acc);
      } else {
        ACC tap = a.use_weights ? (ACC)(w * v) : v;
        if constexpr (!std::is_same<ACC, float>::value)
          if (a.use_weights) tap = (ACC)((uint32_t)w * (uint32_t)v);
        acc = first ? tap : reduce_combine<ACC>(a.reduce, acc, tap);
      }
      first = false;
    }
  }
  const long long oidx = (long long)(g.out_oy + y) * g.out_stride + g.out_ox + x;
  if constexpr (std::is_same<ACC, float>::value)
    store_f(out, g.out_dtype, oidx, acc, g.cvt, g.post_div);
  else
    store_i(out, g.out_dtype, oidx, acc, g.cvt, g.post_div);
}

static int launch_generic(const LocalArgs& a, int acc_is_float, hipStream_t s) {
  dim3 grid(ceil_div(a.g.is_w, 64), ceil_div(a.g.is_h, 4), a.g.batch);
  if (acc_is_float)
    hipLaunchKernelGGL(local_generic_kernel<float>, grid, dim3(64, 4), 0, s, a);
  else
    hipLaunchKernelGGL(local_generic_kernel<int32_t>, grid, dim3(64, 4), 0, s, a);
  return hip_status(hipGetLastError(), "local_generic_kernel launch");
}

static int launch_local_variant(const LocalArgs& f, int acc_is_float, hipStream_t s, int variant,
                                int param) {
  switch (variant) {
    case kVarStream:
      return dispatch_stream_local(f, acc_is_float, param, 0, s);
    case kVarStreamWide:
      return dispatch_stream_local(f, acc_is_float, param, 1, s);
    case kVarRing:
      if (f.g.in_dtype == HIPACC_U8) return dispatch_ring8_local(f, acc_is_float, param, s);
      return acc_is_float ? dispatch_ring_local(f, param, s) : HIPACC_ERR_UNSUPPORTED;
    case kVarTile: {
      if (!acc_is_float) return dispatch_sum_int(f, param, s);
      const int st = dispatch_sum_f32(f, param, s);
      return st == HIPACC_ERR_UNSUPPORTED ? dispatch_sum_u8(f, param, s) : st;
    }
    default:
      return HIPACC_ERR_UNSUPPORTED;
  }
}

// variant kVarAuto: the tuned choice if any, else ring -> stream -> tile ->
// generic (8-bit inputs skip the ring on narrow images, where the streaming
// kernel is faster).
// A forced variant returns HIPACC_ERR_UNSUPPORTED when it has no kernel for
// the operator's shape.
int launch_local(const LocalArgs& a, int acc_is_float, hipStream_t s, int variant, int param) {
  if (dtype_is_vec(a.g.in_dtype))
    return variant == kVarAuto || variant == kVarGeneric ? launch_local_vec4(a, acc_is_float, s)
                                                         : HIPACC_ERR_UNSUPPORTED;
  // The tile and stream kernels implement SUM with weights (or a full domain);
  // domain holes become zero weights, which leaves every finite sum unchanged.
  bool fast = a.reduce == HIPACC_REDUCE_SUM && a.mask_w == a.mask_h;
  if (fast && a.wdev) {
    for (int t = 0; t < a.mask_w * a.mask_h; ++t)
      if (!a.domain[t]) fast = false;
    if (!a.use_weights) fast = false;
  }
  if (variant == kVarGeneric) return launch_generic(a, acc_is_float, s);
  if (!fast) return variant == kVarAuto ? launch_generic(a, acc_is_float, s) : HIPACC_ERR_UNSUPPORTED;
  LocalArgs f = a;
  const int n = a.mask_w * a.mask_h;
  for (int t = 0; t < n; ++t) {
    if (acc_is_float) {
      float w = a.use_weights ? a.w.f[t] : 1.0f;
      f.w.f[t] = a.domain[t] ? w : 0.0f;
    } else {
      int32_t w = a.use_weights ? a.w.i[t] : 1;
      f.w.i[t] = a.domain[t] ? w : 0;
    }
  }
  f.use_weights = 1;
  if (variant != kVarAuto) return launch_local_variant(f, acc_is_float, s, variant, param);
  TuneKey k{kTuneLocal, a.mask_w, a.mask_h, a.g.in_dtype, a.g.out_dtype, acc_is_float,
            a.g.boundary, a.fma, a.wdev != nullptr, tune_size_class(a.g.is_w, a.g.is_h, a.g.batch)};
  TuneVal v;
  if (tune_lookup(k, v)) {
    if (v.variant == kVarGeneric) return launch_generic(a, acc_is_float, s);
    const int st = launch_local_variant(f, acc_is_float, s, v.variant, v.param);
    if (st != HIPACC_ERR_UNSUPPORTED) return st;
  }
  int st = HIPACC_ERR_UNSUPPORTED;
  if (a.g.in_dtype != HIPACC_U8 || ring8_pays(f))
    st = launch_local_variant(f, acc_is_float, s, kVarRing, 0);
  if (st == HIPACC_ERR_UNSUPPORTED) st = launch_local_variant(f, acc_is_float, s, kVarStream, 0);
  if (st == HIPACC_ERR_UNSUPPORTED) st = launch_local_variant(f, acc_is_float, s, kVarTile, 0);
  if (st == HIPACC_ERR_UNSUPPORTED) st = launch_generic(a, acc_is_float, s);
  return st;
}

}  // namespace hh
