// hipacc_hip_device.hpp -- the DSL's kernel-side API on the device, for
// kernels that are not one of the library's operators.
//
// A DSL Kernel class whose kernel() the library does not provide as an
// operator (arbitrary per-pixel code: Sharpen's `2 * Gray() - Blur()`, the
// bilateral sample's iterate() body, Harris' response, ...) is compiled as
// it is, as HIP device code, against the classes here: the emitter writes a
// struct deriving from hipacc_dev::Kernel<T> with the class's members as
// hipacc_dev::Accessor / Mask / Domain values and its kernel() body
// unchanged, and launches it with hipaccLaunchUser over the iteration space.
// The semantics are the DSL's (dsl/kernel.hpp:241-315, dsl/image.hpp:
// 536-708, dsl/mask.hpp): convolve() visits every mask position row by row,
// reduce() and iterate() the domain's enabled positions, the first value
// seeds the reduction; accessors read through their boundary handling
// relative to the iteration-space pixel; Mask()/Domain() give the current
// position's coefficient and offset.  Compile with hipcc (-ffp-contract=off
// for the DSL host emulation's rounding, the default contraction for the
// reference CUDA backend's).
#ifndef HIPACC_HIP_DEVICE_HPP
#define HIPACC_HIP_DEVICE_HPP

#include <hip/hip_runtime.h>

#include <algorithm>
#include <memory>
#include <type_traits>
#include <vector>

#include "../csrc/hh_pyramid.hpp"
#include "hipacc_hip.hpp"
#include "hipacc_hip_vec.hpp"

namespace hipacc_dev {

enum class Reduce : int { SUM = 0, MIN, MAX, PROD, MEDIAN };

template <typename T> __device__ __forceinline__ T dsl_min(T a, T b) { return a < b ? a : b; }
template <typename T> __device__ __forceinline__ T dsl_max(T a, T b) { return a > b ? a : b; }
// 4-wide pixels: componentwise (dsl/math_functions.hpp:345-392).
template <typename E>
__device__ __forceinline__ HIP_vector_type<E, 4> dsl_min(HIP_vector_type<E, 4> a,
                                                         HIP_vector_type<E, 4> b) {
  return hipacc::math::min(a, b);
}
template <typename E>
__device__ __forceinline__ HIP_vector_type<E, 4> dsl_max(HIP_vector_type<E, 4> a,
                                                         HIP_vector_type<E, 4> b) {
  return hipacc::math::max(a, b);
}

// An accessor: the image window it reads, its boundary handling, its
// interpolation, and the iteration-space pixel (px, py) the kernel is at.
// INTERP: the accessor may be bound to an interpolating window (the code of
// the interpolation modes is compiled in only then).
template <typename T, bool INTERP = false> struct Accessor {
  const T* img = nullptr;
  int stride = 0, img_w = 0, img_h = 0;
  int lo_x = 0, lo_y = 0, w = 0, h = 0;
  int mode = HIPACC_BOUNDARY_UNDEFINED;
  T cval{};
  int interp = 0;                 // 0 or HIPACC_INTERP_NN ... L3
  float isx = 1.0f, isy = 1.0f;   // window extent / iteration-space extent
  int px = 0, py = 0;

  // Pixel (x, y) of the accessor's window (window coordinates), through the
  // boundary handling.  UNDEFINED reads are clamped to the allocation.
  __device__ __forceinline__ T at(int x, int y) const {
    int ax = lo_x + x, ay = lo_y + y;
    if (mode == HIPACC_BOUNDARY_CONSTANT) {
      if (x < 0 || x >= w || y < 0 || y >= h) return cval;
    } else {
      ax = hh::bh_map(mode, ax, lo_x, lo_x + w, img_w);
      ay = hh::bh_map(mode, ay, lo_y, lo_y + h, img_h);
    }
    return img[(long long)ay * stride + ax];
  }
  // Iteration-space pixel (x, y) through the interpolation, if any
  // (dsl/image.hpp:390-532): mapped into the window (cell-centred, scale
  // window / iteration-space extent); NN returns the pixel there, the other
  // modes their float value converted to T with a C cast.
  __device__ __forceinline__ T read(int x, int y) const {
    if (!INTERP || interp == 0) return at(x, y);
    const float mx = hh::interp_map(isx, x, lo_x), my = hh::interp_map(isy, y, lo_y);
    if (interp == HIPACC_INTERP_NN) return at((int)mx - lo_x, (int)my - lo_y);
    if constexpr (std::is_arithmetic<T>::value) {
      auto rd = [&](int ax, int ay) { return (float)at(ax - lo_x, ay - lo_y); };
      switch (interp) {
        case HIPACC_INTERP_LF: return (T)hh::interp_sample<HIPACC_INTERP_LF>(rd, mx, my);
        case HIPACC_INTERP_B5: return (T)hh::interp_sample<HIPACC_INTERP_B5>(rd, mx, my);
        case HIPACC_INTERP_CF: return (T)hh::interp_sample<HIPACC_INTERP_CF>(rd, mx, my);
        default: return (T)hh::interp_sample<HIPACC_INTERP_L3>(rd, mx, my);
      }
    } else {
      return at(x, y);    // 4-wide pixels: the host allows NN only
    }
  }
  __device__ __forceinline__ T operator()() const { return read(px, py); }
  __device__ __forceinline__ T operator()(int xf, int yf) const { return read(px + xf, py + yf); }
  template <typename M> __device__ __forceinline__ T operator()(const M& m) const {
    return read(px + m.x(), py + m.y());
  }
  __device__ __forceinline__ T pixel_at(int x, int y) const {
    return img[(long long)(lo_y + y) * stride + lo_x + x];
  }
  __device__ __forceinline__ int x() const { return px; }
  __device__ __forceinline__ int y() const { return py; }
};

// A read of a table the kernel never writes (mask coefficients, domain
// flags) through the constant address space: at a wave-uniform index (the
// position convolve / reduce / iterate is at) it becomes a scalar load.
template <typename T> __device__ __forceinline__ T table_at(const T* p, int i) {
  return ((const __attribute__((address_space(4))) T*)p)[i];
}

// A domain: enabled positions of a size_x x size_y window; (cx, cy) is the
// position reduce() / iterate() is at.
struct Domain {
  const uint8_t* on = nullptr;
  int sx = 1, sy = 1;
  int cx = 0, cy = 0;
  __device__ __forceinline__ int x() const { return cx - sx / 2; }
  __device__ __forceinline__ int y() const { return cy - sy / 2; }
  __device__ __forceinline__ int size_x() const { return sx; }
  __device__ __forceinline__ int size_y() const { return sy; }
};

// A mask: coefficients of a size_x x size_y window (device memory, row
// major); (cx, cy) is the position convolve() is at.
template <typename T> struct Mask {
  const T* coef = nullptr;
  int sx = 1, sy = 1;
  int cx = 0, cy = 0;
  __device__ __forceinline__ int x() const { return cx - sx / 2; }
  __device__ __forceinline__ int y() const { return cy - sy / 2; }
  __device__ __forceinline__ int size_x() const { return sx; }
  __device__ __forceinline__ int size_y() const { return sy; }
  __device__ __forceinline__ T operator()() const { return table_at(coef, cy * sx + cx); }
  __device__ __forceinline__ T operator()(int xf, int yf) const {
    return table_at(coef, (yf + sy / 2) * sx + xf + sx / 2);
  }
  __device__ __forceinline__ T operator()(const Domain& d) const {
    return table_at(coef, (d.y() + d.sy / 2) * sx + d.x() + d.sx / 2);
  }
};

// The kernel base: output() and the iteration-space pixel, and the DSL's
// convolve / reduce / iterate.
template <typename TO> struct Kernel {
  TO out_{};
  int it_x = 0, it_y = 0;
  __device__ __forceinline__ TO& output() { return out_; }
  __device__ __forceinline__ int x() const { return it_x; }
  __device__ __forceinline__ int y() const { return it_y; }

  template <typename V> __device__ __forceinline__ static V combine(Reduce mode, V result, V v) {
    switch (mode) {
      case Reduce::SUM: result += v; break;
      case Reduce::MIN: result = dsl_min(v, result); break;
      case Reduce::MAX: result = dsl_max(v, result); break;
      case Reduce::PROD: result *= v; break;
      default: break;
    }
    return result;
  }
  template <typename M, typename F>
  __device__ __forceinline__ auto convolve(Mask<M>& mask, Reduce mode, const F& fun) -> decltype(fun()) {
    mask.cx = mask.cy = 0;
    auto result = fun();
    for (int i = 1; i < mask.sx * mask.sy; ++i) {
      mask.cy = i / mask.sx;
      mask.cx = i - mask.cy * mask.sx;
      result = combine(mode, result, fun());
    }
    return result;
  }
  template <typename F>
  __device__ __forceinline__ auto reduce(Domain& dom, Reduce mode, const F& fun) -> decltype(fun()) {
    int i = 0;
    const int n = dom.sx * dom.sy;
    while (i < n && !table_at(dom.on, i)) ++i;
    dom.cy = i / dom.sx;
    dom.cx = i - dom.cy * dom.sx;
    auto result = fun();
    for (++i; i < n; ++i) {
      if (!table_at(dom.on, i)) continue;
      dom.cy = i / dom.sx;
      dom.cx = i - dom.cy * dom.sx;
      result = combine(mode, result, fun());
    }
    return result;
  }
  template <typename F> __device__ __forceinline__ void iterate(Domain& dom, const F& fun) {
    for (int i = 0; i < dom.sx * dom.sy; ++i) {
      if (!table_at(dom.on, i)) continue;
      dom.cy = i / dom.sx;
      dom.cx = i - dom.cy * dom.sx;
      fun();
    }
  }
};

// One output pixel per thread, 64 x 4 blocks.  F::set_pos(x, y) points the
// kernel and its accessors at iteration-space pixel (x, y).
template <typename F, typename TO>
__global__ __launch_bounds__(256) void user_kernel(F f, TO* out, int stride, int ox, int oy,
                                                   int w, int h) {
  const int x = blockIdx.x * 64 + threadIdx.x, y = blockIdx.y * 4 + threadIdx.y;
  if (x >= w || y >= h) return;
  F k = f;
  k.set_pos(x, y);
  k.kernel();
  out[(long long)(oy + y) * stride + ox + x] = k.output();
}

// The binning side of a Kernel<T, TB> class: binning(x, y, pixel) sets
// bin(idx) = value, reduce(l, r) combines values (dsl/kernel.hpp:163-209).
template <typename TB> struct Binning {
  unsigned bidx_ = 0xffffffffu;
  TB bval_{};
  unsigned nbins_ = 0;
  __device__ __forceinline__ TB& bin(unsigned idx) {
    bidx_ = idx;
    return bval_;
  }
  __device__ __forceinline__ unsigned num_bins() const { return nbins_; }
};

// Combines v into the LDS bin *p with the user's reduce(), as one atomic
// update of the 32-bit word.  RED names the reduce() when it is one of the
// DSL's + / min / max forms (HIPACC_REDUCE_*; -1 otherwise): then the LDS
// atomic of that operation does it in one step; any other reduce() is a
// compare-and-swap loop (until no other lane changed the word in between).
template <int RED, typename F, typename TB>
__device__ __forceinline__ void combine_at(const F& k, TB* p, TB v) {
  static_assert(sizeof(TB) == 4, "bins of 32-bit values");
  constexpr bool integral = std::is_integral<TB>::value;
  if constexpr (RED == HIPACC_REDUCE_SUM && (integral || std::is_same<TB, float>::value)) {
    atomicAdd(p, v);
  } else if constexpr (RED == HIPACC_REDUCE_MIN && integral) {
    atomicMin(p, v);
  } else if constexpr (RED == HIPACC_REDUCE_MAX && integral) {
    atomicMax(p, v);
  } else {
    unsigned* w = reinterpret_cast<unsigned*>(p);
    unsigned old = *w, seen;
    do {
      seen = old;
      const TB nv = k.reduce(__builtin_bit_cast(TB, seen), v);
      old = atomicCAS(w, seen, __builtin_bit_cast(unsigned, nv));
    } while (old != seen);
  }
}

// Bins of the output image, 32-bit values: each block combines its pixels'
// values into bins in LDS (starting from TB{}, as the DSL's binned_data()
// does) and stores them as one row of `rows`; user_binning_fold then folds
// the rows bin by bin in block order.  reduce() must be associative and
// commutative, as any parallel binning assumes.
template <int RED, typename F, typename T, typename TB>
__global__ __launch_bounds__(256) void user_binning_kernel(F f, const T* img, int stride, int ox,
                                                           int oy, int w, int h, TB* rows,
                                                           unsigned nbins) {
  extern __shared__ __align__(16) unsigned char user_bins_lds[];
  TB* lb = reinterpret_cast<TB*>(user_bins_lds);
  for (unsigned b = threadIdx.x; b < nbins; b += 256) lb[b] = TB{};
  __syncthreads();
  F k = f;
  k.nbins_ = nbins;
  const long long n = (long long)w * h;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (long long)gridDim.x * 256) {
    const int y = (int)(i / w), x = (int)(i - (long long)y * w);
    k.bidx_ = 0xffffffffu;
    k.binning((unsigned)x, (unsigned)y, img[(long long)(oy + y) * stride + ox + x]);
    if (k.bidx_ < nbins) combine_at<RED>(k, lb + k.bidx_, k.bval_);
  }
  __syncthreads();
  for (unsigned b = threadIdx.x; b < nbins; b += 256) rows[(long long)blockIdx.x * nbins + b] = lb[b];
}

// Bins of values wider than 32 bits (no atomic update of such a value):
// one wave per block; the wave's lanes that have a bin update the block's
// LDS bins one lane at a time, so no two updates of a bin overlap; each
// block stores its bins as one row of `rows`, and user_binning_fold then
// folds the rows bin by bin in block order.
template <typename F, typename T, typename TB>
__global__ __launch_bounds__(64) void user_binning_wide_kernel(F f, const T* img, int stride,
                                                               int ox, int oy, int w, int h,
                                                               TB* rows, unsigned nbins) {
  extern __shared__ __align__(16) unsigned char user_bins_lds[];
  TB* lb = reinterpret_cast<TB*>(user_bins_lds);
  const int lane = threadIdx.x;
  for (unsigned b = lane; b < nbins; b += 64) lb[b] = TB{};
  __syncthreads();
  F k = f;
  k.nbins_ = nbins;
  const long long n = (long long)w * h;
  for (long long base = (long long)blockIdx.x * 64; base < n; base += (long long)gridDim.x * 64) {
    const long long i = base + lane;
    bool have = false;
    if (i < n) {
      const int y = (int)(i / w), x = (int)(i - (long long)y * w);
      k.bidx_ = 0xffffffffu;
      k.binning((unsigned)x, (unsigned)y, img[(long long)(oy + y) * stride + ox + x]);
      have = k.bidx_ < nbins;
    }
    unsigned long long pending = __ballot(have);
    while (pending) {
      const int l = __ffsll((long long)pending) - 1;
      pending &= pending - 1;
      if (lane == l) lb[k.bidx_] = k.reduce(lb[k.bidx_], k.bval_);
    }
  }
  __syncthreads();
  for (unsigned b = lane; b < nbins; b += 64) rows[(long long)blockIdx.x * nbins + b] = lb[b];
}

template <typename F, typename TB>
__global__ __launch_bounds__(256) void user_binning_fold(F f, const TB* rows, int nrows,
                                                         unsigned nbins, TB* bins) {
  const unsigned b = blockIdx.x * 256 + threadIdx.x;
  if (b >= nbins) return;
  const F k = f;
  TB acc = rows[b];
  for (int r = 1; r < nrows; ++r) acc = k.reduce(acc, rows[(long long)r * nbins + b]);
  bins[b] = acc;
}

// A user reduction reduce(l, r) over an image window: per-block trees in
// LDS, then one block folds the per-block results, both in a fixed order.
template <typename F, typename T>
__global__ __launch_bounds__(256) void user_reduce_kernel(F f, const T* img, int stride, int ox,
                                                          int oy, int w, int h, T* partial) {
  __shared__ T red[256];
  const F k = f;
  const long long n = (long long)w * h;
  const long long per = (n + gridDim.x - 1) / gridDim.x;
  const long long i0 = per * blockIdx.x, i1 = i0 + per < n ? i0 + per : n;
  bool have = false;
  T acc{};
  for (long long i = i0 + threadIdx.x; i < i1; i += 256) {
    const int y = (int)(i / w), x = (int)(i - (long long)y * w);
    const T v = img[(long long)(oy + y) * stride + ox + x];
    acc = have ? k.reduce(acc, v) : v;
    have = true;
  }
  __shared__ int valid[256];
  red[threadIdx.x] = acc;
  valid[threadIdx.x] = have;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s && valid[threadIdx.x + s]) {
      red[threadIdx.x] = valid[threadIdx.x] ? k.reduce(red[threadIdx.x], red[threadIdx.x + s])
                                            : red[threadIdx.x + s];
      valid[threadIdx.x] = 1;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

template <typename F, typename T>
__global__ __launch_bounds__(64) void user_reduce_final(F f, T* partial, int nparts) {
  if (threadIdx.x != 0) return;
  const F k = f;
  T acc = partial[0];
  for (int i = 1; i < nparts; ++i) acc = k.reduce(acc, partial[i]);
  partial[0] = acc;
}

}  // namespace hipacc_dev

// ---- host side ---------------------------------------------------------------
// The device view of a runtime accessor with the boundary handling of its
// BoundaryCondition.
// With an interpolation (HIPACC_INTERP_NN ... L3; 0 for none), `is` is the
// iteration space the accessor is read over: its extent against the
// accessor's sets the mapping.
template <bool INTERP = false, typename T>
hipacc_dev::Accessor<T, INTERP> hipaccDevAccessor(const HipaccAccessor<T>& acc, HipaccBoundary mode,
                                          double const_val = 0.0, int interp = 0,
                                          const HipaccAccessorBase* is = nullptr) {
  const hipacc_view v = hipaccView(acc);
  hipacc_dev::Accessor<T, INTERP> d;
  d.img = static_cast<const T*>(v.data);
  d.stride = v.stride;
  d.img_w = v.img_width;
  d.img_h = v.img_height;
  d.lo_x = v.offset_x;
  d.lo_y = v.offset_y;
  d.w = v.width;
  d.h = v.height;
  d.mode = (int)mode;
  d.cval = (T)const_val;
  if (interp) {
    if (!INTERP || (!std::is_arithmetic<T>::value && interp != HIPACC_INTERP_NN) || !is ||
        is->width <= 0 || is->height <= 0) {
      hipaccRuntimeLogTrivial(hipaccRuntimeLogLevel::ERROR,
                              "ERROR: hipaccDevAccessor: interpolation needs an accessor "
                              "made with INTERP, the iteration space, and a scalar pixel "
                              "type unless it is NN");
    } else {
      d.interp = interp;
      d.isx = (float)v.width / (float)is->width;
      d.isy = (float)v.height / (float)is->height;
    }
  }
  return d;
}

// Device copies of mask coefficients and domain bitmaps, kept for the life
// of the process (a mask's table is written once and read by every launch).
inline const void* hipaccDevTable(const void* host, size_t bytes) {
  static std::vector<std::shared_ptr<void>> tables;
  void* p = nullptr;
  int32_t stride = 0;
  if (hipaccCheck(hipacc_hip_alloc2d(&p, &stride, (int32_t)bytes, 1, 1, 0), "hipaccDevTable"))
    return nullptr;
  tables.emplace_back(p, [](void* q) { (void)hipacc_hip_free(q); });
  hipaccCheck(hipacc_hip_upload2d(p, stride, host, (int32_t)bytes, 1, 1, nullptr),
              "hipaccDevTable");
  return p;
}

template <typename M> hipacc_dev::Mask<M> hipaccDevMask(const HipaccMask<M>& m) {
  hipacc_dev::Mask<M> d;
  d.coef = static_cast<const M*>(hipaccDevTable(m.coef.data(), m.coef.size() * sizeof(M)));
  d.sx = m.size_x;
  d.sy = m.size_y;
  return d;
}

template <typename M> hipacc_dev::Domain hipaccDevDomain(const HipaccMask<M>& m) {
  hipacc_dev::Domain d;
  d.on = static_cast<const uint8_t*>(hipaccDevTable(m.domain.data(), m.domain.size()));
  d.sx = m.size_x;
  d.sy = m.size_y;
  return d;
}

// Times the launches made here (not in the library) when the calling
// thread's timing is on, as the library's entry points time theirs: events
// around the launches on their stream, a synchronisation, and the elapsed
// time handed to hipacc_last_kernel_timing().
class HipaccLaunchTimer {
  hipStream_t s_;
  hipEvent_t e0_ = nullptr, e1_ = nullptr;

 public:
  explicit HipaccLaunchTimer(hipStream_t s) : s_(s) {
    if (hipacc_hip_get_timing() && hipEventCreate(&e0_) == hipSuccess &&
        hipEventCreate(&e1_) == hipSuccess)
      (void)hipEventRecord(e0_, s_);
  }
  void finish(int status) {
    if (e0_ && e1_ && status == HIPACC_OK) {
      float ms = 0.0f;
      if (hipEventRecord(e1_, s_) == hipSuccess && hipEventSynchronize(e1_) == hipSuccess &&
          hipEventElapsedTime(&ms, e0_, e1_) == hipSuccess)
        (void)hipacc_hip_set_last_kernel_timing(ms);
    }
  }
  ~HipaccLaunchTimer() {
    if (e0_) (void)hipEventDestroy(e0_);
    if (e1_) (void)hipEventDestroy(e1_);
  }
  HipaccLaunchTimer(const HipaccLaunchTimer&) = delete;
  HipaccLaunchTimer& operator=(const HipaccLaunchTimer&) = delete;
};

// Launches a user kernel over the iteration space (a DSL image: one plane).
template <typename F, typename TO>
int hipaccLaunchUser(const F& f, const HipaccAccessor<TO>& is,
                     const HipaccExecutionParameterHip& ep = nullptr) {
  const hipacc_view v = hipaccView(is);
  dim3 grid((v.width + 63) / 64, (v.height + 3) / 4, 1);
  hipStream_t s = static_cast<hipStream_t>(hipaccStreamOf(ep));
  if (ep) ep->pre_kernel();
  HipaccLaunchTimer timer(s);
  hipLaunchKernelGGL((hipacc_dev::user_kernel<F, TO>), grid, dim3(64, 4), 0, s, f,
                     static_cast<TO*>(v.data), v.stride, v.offset_x, v.offset_y, v.width,
                     v.height);
  const hipError_t e = hipGetLastError();
  const int st = hipaccCheck(e == hipSuccess ? HIPACC_OK : HIPACC_ERR_HIP, "hipaccLaunchUser");
  timer.finish(st);
  if (ep) ep->post_kernel();
  return st;
}

// ---------------------------------------------------------------------------
// Launching a kernel of the program's own in the reference's generated-code
// form (runtime/hipacc_base.hpp:87-101, runtime/hipacc_base_standalone.hpp:
// 41-58, runtime/hipacc_cu_standalone.hpp:66-110,291-329): the launch
// geometry of an iteration space, the grid for a block shape, the block
// indices between which a block needs no border handling, and a launch on an
// execution parameter's stream, timed when timing is on.
// ---------------------------------------------------------------------------
struct hipacc_launch_info {
  // size_x / size_y: half the largest window read with boundary handling
  // (0 for none); the iteration space's extent and offset; pixels per
  // thread (rows) and SIMD width (columns).
  hipacc_launch_info(int size_x_, int size_y_, int is_width_, int is_height_, int offset_x_,
                     int offset_y_, int pixels_per_thread_, int simd_width_)
      : size_x(size_x_), size_y(size_y_), is_width(is_width_), is_height(is_height_),
        offset_x(offset_x_), offset_y(offset_y_), pixels_per_thread(pixels_per_thread_),
        simd_width(simd_width_) {}
  hipacc_launch_info(int size_x_, int size_y_, const HipaccAccessorBase& is,
                     int pixels_per_thread_, int simd_width_)
      : hipacc_launch_info(size_x_, size_y_, is.width, is.height, is.offset_x, is.offset_y,
                           pixels_per_thread_, simd_width_) {}
  int size_x, size_y;
  int is_width, is_height;
  int offset_x, offset_y;
  int pixels_per_thread, simd_width;
  // set by hipaccPrepareKernelLaunch
  int bh_start_left = 0, bh_start_right = 0;
  int bh_start_top = 0, bh_start_bottom = 0;
  int bh_fall_back = 0;
};

// Blocks covering the iteration space: block.x * simd_width columns and
// block.y * pixels_per_thread rows each (the columns from 0, the offset
// included, as the generated kernels index them).
inline dim3 hipaccCalcGridFromBlock(const hipacc_launch_info& info, const dim3& block) {
  const int bw = (int)block.x * info.simd_width, bh = (int)block.y * info.pixels_per_thread;
  return dim3((unsigned)((info.is_width + info.offset_x + bw - 1) / bw),
              (unsigned)((info.is_height + bh - 1) / bh));
}

// Block columns [bh_start_left, bh_start_right) and rows [bh_start_top,
// bh_start_bottom) need no border handling: their pixels' windows stay
// inside the image.  The left bound is the first block whose first column
// is past offset_x + size_x; the right bound the blocks that end before the
// last size_x columns; the bottom bound keeps 2 * size_y rows of blocks
// back, room for the rows a shared-memory variant stages below a block.  With
// fewer than two interior blocks either way every block takes the border
// path (bh_fall_back).
inline void hipaccPrepareKernelLaunch(hipacc_launch_info& info, const dim3& block) {
  const int bw = (int)block.x * info.simd_width;
  const int bh = (int)block.y * info.pixels_per_thread;
  auto div_up = [](int a, int b) { return a <= 0 ? -((-a) / b) : (a + b - 1) / b; };
  auto div_down = [](int a, int b) { return a >= 0 ? a / b : -((-a + b - 1) / b); };
  if (info.size_x > 0) {
    info.bh_start_left = div_up(info.offset_x + info.size_x, bw);
    info.bh_start_right = div_down(info.offset_x + info.is_width - info.size_x, bw);
  } else {
    info.bh_start_left = 0;
    info.bh_start_right = div_down(info.offset_x + info.is_width, bw);
  }
  if (info.size_y > 0) {
    const int staged = div_up(2 * info.size_y, (int)block.y);
    info.bh_start_top = div_up(info.size_y, bh);
    info.bh_start_bottom = div_down(info.is_height - staged * (int)block.y, bh);
  } else {
    info.bh_start_top = 0;
    info.bh_start_bottom = div_down(info.is_height, bh);
  }
  info.bh_fall_back = (info.bh_start_right - info.bh_start_left) > 1 &&
                              (info.bh_start_bottom - info.bh_start_top) > 1
                          ? 0
                          : 1;
}

// hipaccLaunchKernel(kernel, grid, block, ep, print_timing, shared_memory,
// args...): the launch of a __global__ function on the execution
// parameter's stream; with print_timing (or the calling thread's timing
// switch on) the launch is timed for hipacc_last_kernel_timing().
template <typename... P, typename... A>
int hipaccLaunchKernel(void (*kernel)(P...), const dim3& grid, const dim3& block,
                       const HipaccExecutionParameterHip& ep, bool print_timing,
                       size_t shared_memory, A&&... args) {
  hipStream_t s = static_cast<hipStream_t>(hipaccStreamOf(ep));
  const int was = hipacc_hip_get_timing();
  if (print_timing) hipacc_hip_set_timing(1);
  if (ep) ep->pre_kernel();
  int st;
  {
    HipaccLaunchTimer timer(s);
    hipLaunchKernelGGL(kernel, grid, block, (unsigned)shared_memory, s, std::forward<A>(args)...);
    const hipError_t e = hipGetLastError();
    st = hipaccCheck(e == hipSuccess ? HIPACC_OK : HIPACC_ERR_HIP, "hipaccLaunchKernel");
    timer.finish(st);
  }
  if (ep) ep->post_kernel();
  if (print_timing) hipacc_hip_set_timing(was);
  return st;
}

// binned_data(num_bins) of a Kernel<T, TB> class over the image its
// iteration space wrote: a new TB[num_bins] the caller owns.  Blocking.
// RED: HIPACC_REDUCE_SUM / MIN / MAX when reduce() is that operation (one
// LDS atomic per pixel), -1 for any other reduce() (compare-and-swap).
template <typename F, typename TB, int RED = -1, typename T>
TB* hipaccBinUser(const F& f, const HipaccAccessor<T>& is, unsigned num_bins,
                  const HipaccExecutionParameterHip& ep = nullptr) {
  TB* result = new TB[num_bins]();
  const hipacc_view v = hipaccView(is);
  const size_t lds = (size_t)num_bins * sizeof(TB);
  if (num_bins == 0 || lds > 65536) {
    hipaccRuntimeLogTrivial(hipaccRuntimeLogLevel::ERROR,
                            "ERROR: hipaccBinUser: " + std::to_string(num_bins) +
                                " bins do not fit the 64 KiB of a block's bins");
    return result;
  }
  constexpr bool narrow = sizeof(TB) == 4;
  const long long n = (long long)v.width * v.height;
  const int blocks = (int)std::min<long long>(std::max<long long>((n + 4095) / 4096, 1),
                                              narrow ? 1024 : 512);
  // the bins, then one row of bins per block
  const size_t words = (lds + 3) / 4 + (size_t)blocks * lds / 4;
  void* bins = nullptr;
  int32_t stride = 0;
  if (hipaccCheck(hipacc_hip_alloc2d(&bins, &stride, (int32_t)words, 1, 4, 0), "hipaccBinUser"))
    return result;
  hipStream_t s = static_cast<hipStream_t>(hipaccStreamOf(ep));
  if (ep) ep->pre_kernel();
  HipaccLaunchTimer timer(s);
  TB* out = static_cast<TB*>(bins);
  TB* rows = reinterpret_cast<TB*>(static_cast<char*>(bins) + (lds + 3) / 4 * 4);
  if constexpr (narrow) {
    hipLaunchKernelGGL((hipacc_dev::user_binning_kernel<RED, F, T, TB>), dim3(blocks), dim3(256),
                       lds, s, f, static_cast<const T*>(v.data), v.stride, v.offset_x,
                       v.offset_y, v.width, v.height, rows, num_bins);
  } else {
    hipLaunchKernelGGL((hipacc_dev::user_binning_wide_kernel<F, T, TB>), dim3(blocks), dim3(64),
                       lds, s, f, static_cast<const T*>(v.data), v.stride, v.offset_x,
                       v.offset_y, v.width, v.height, rows, num_bins);
  }
  hipLaunchKernelGGL((hipacc_dev::user_binning_fold<F, TB>), dim3((num_bins + 255) / 256),
                     dim3(256), 0, s, f, rows, blocks, num_bins, out);
  const hipError_t e = hipGetLastError();
  const int st = hipaccCheck(e == hipSuccess ? HIPACC_OK : HIPACC_ERR_HIP, "hipaccBinUser");
  timer.finish(st);
  if (!st)
    hipaccCheck(hipacc_hip_download2d(result, bins, (int32_t)lds, (int32_t)lds, 1, 1, s),
                "hipaccBinUser");
  if (ep) ep->post_kernel();
  (void)hipacc_hip_free(bins);
  return result;
}

// reduced_data() of a Kernel class with a reduce(l, r) the library has no
// operator for, over the image its iteration space wrote.  Blocking.
template <typename F, typename T>
T hipaccReduceUser(const F& f, const HipaccAccessor<T>& is,
                   const HipaccExecutionParameterHip& ep = nullptr) {
  T result{};
  const hipacc_view v = hipaccView(is);
  constexpr int kBlocks = 256;
  void* part = nullptr;
  int32_t stride = 0;
  if (hipaccCheck(hipacc_hip_alloc2d(&part, &stride, kBlocks, 1, (int32_t)sizeof(T), 0),
                  "hipaccReduceUser"))
    return result;
  hipStream_t s = static_cast<hipStream_t>(hipaccStreamOf(ep));
  const long long n = (long long)v.width * v.height;
  const int blocks = (int)std::min<long long>(std::max<long long>((n + 255) / 256, 1), kBlocks);
  if (ep) ep->pre_kernel();
  HipaccLaunchTimer timer(s);
  hipLaunchKernelGGL((hipacc_dev::user_reduce_kernel<F, T>), dim3(blocks), dim3(256), 0, s, f,
                     static_cast<const T*>(v.data), v.stride, v.offset_x, v.offset_y, v.width,
                     v.height, static_cast<T*>(part));
  hipLaunchKernelGGL((hipacc_dev::user_reduce_final<F, T>), dim3(1), dim3(64), 0, s, f,
                     static_cast<T*>(part), blocks);
  const hipError_t e = hipGetLastError();
  const int st = hipaccCheck(e == hipSuccess ? HIPACC_OK : HIPACC_ERR_HIP, "hipaccReduceUser");
  timer.finish(st);
  if (!st)
    hipaccCheck(hipacc_hip_download2d(&result, part, 1, 1, 1, (int32_t)sizeof(T), s),
                "hipaccReduceUser");
  if (ep) ep->post_kernel();
  (void)hipacc_hip_free(part);
  return result;
}

#endif  // HIPACC_HIP_DEVICE_HPP
