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
#include <vector>

#include "../csrc/hh_common.hpp"
#include "hipacc_hip.hpp"

namespace hipacc_dev {

enum class Reduce : int { SUM = 0, MIN, MAX, PROD, MEDIAN };

template <typename T> __device__ __forceinline__ T dsl_min(T a, T b) { return a < b ? a : b; }
template <typename T> __device__ __forceinline__ T dsl_max(T a, T b) { return a > b ? a : b; }

// An accessor: the image window it reads, its boundary handling, and the
// iteration-space pixel (px, py) the kernel is at.
template <typename T> struct Accessor {
  const T* img = nullptr;
  int stride = 0, img_w = 0, img_h = 0;
  int lo_x = 0, lo_y = 0, w = 0, h = 0;
  int mode = HIPACC_BOUNDARY_UNDEFINED;
  T cval{};
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
  __device__ __forceinline__ T operator()() const { return at(px, py); }
  __device__ __forceinline__ T operator()(int xf, int yf) const { return at(px + xf, py + yf); }
  template <typename M> __device__ __forceinline__ T operator()(const M& m) const {
    return at(px + m.x(), py + m.y());
  }
  __device__ __forceinline__ T pixel_at(int x, int y) const {
    return img[(long long)(lo_y + y) * stride + lo_x + x];
  }
  __device__ __forceinline__ int x() const { return px; }
  __device__ __forceinline__ int y() const { return py; }
};

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
  __device__ __forceinline__ T operator()() const { return coef[cy * sx + cx]; }
  __device__ __forceinline__ T operator()(int xf, int yf) const {
    return coef[(yf + sy / 2) * sx + xf + sx / 2];
  }
  __device__ __forceinline__ T operator()(const Domain& d) const {
    return coef[(d.y() + d.sy / 2) * sx + d.x() + d.sx / 2];
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
    while (i < n && !dom.on[i]) ++i;
    dom.cy = i / dom.sx;
    dom.cx = i - dom.cy * dom.sx;
    auto result = fun();
    for (++i; i < n; ++i) {
      if (!dom.on[i]) continue;
      dom.cy = i / dom.sx;
      dom.cx = i - dom.cy * dom.sx;
      result = combine(mode, result, fun());
    }
    return result;
  }
  template <typename F> __device__ __forceinline__ void iterate(Domain& dom, const F& fun) {
    for (int i = 0; i < dom.sx * dom.sy; ++i) {
      if (!dom.on[i]) continue;
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

// Combines v into *p with the user's reduce(), as one atomic update of the
// 32-bit word (compare-and-swap until no other thread changed it between).
template <typename F, typename TB>
__device__ __forceinline__ void combine_at(const F& k, TB* p, TB v) {
  static_assert(sizeof(TB) == 4, "bins of 32-bit values");
  unsigned* w = reinterpret_cast<unsigned*>(p);
  unsigned old = *w, seen;
  do {
    seen = old;
    const TB nv = k.reduce(__builtin_bit_cast(TB, seen), v);
    old = atomicCAS(w, seen, __builtin_bit_cast(unsigned, nv));
  } while (old != seen);
}

// Bins of the output image: each block combines its pixels' values into
// bins in LDS, then its bins into the global bins (zero before the launch,
// as the DSL's binned_data() starts from TB{}).  reduce() must be
// associative and commutative, as any parallel binning assumes.
template <typename F, typename T, typename TB>
__global__ __launch_bounds__(256) void user_binning_kernel(F f, const T* img, int stride, int ox,
                                                           int oy, int w, int h, TB* bins,
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
    if (k.bidx_ < nbins) combine_at(k, lb + k.bidx_, k.bval_);
  }
  __syncthreads();
  for (unsigned b = threadIdx.x; b < nbins; b += 256)
    if (lb[b] != TB{}) combine_at(k, bins + b, lb[b]);
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
template <typename T>
hipacc_dev::Accessor<T> hipaccDevAccessor(const HipaccAccessor<T>& acc, HipaccBoundary mode,
                                          double const_val = 0.0) {
  const hipacc_view v = hipaccView(acc);
  hipacc_dev::Accessor<T> d;
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

// Launches a user kernel over the iteration space (a DSL image: one plane).
template <typename F, typename TO>
int hipaccLaunchUser(const F& f, const HipaccAccessor<TO>& is,
                     const HipaccExecutionParameterHip& ep = nullptr) {
  const hipacc_view v = hipaccView(is);
  dim3 grid((v.width + 63) / 64, (v.height + 3) / 4, 1);
  if (ep) ep->pre_kernel();
  hipLaunchKernelGGL((hipacc_dev::user_kernel<F, TO>), grid, dim3(64, 4), 0,
                     static_cast<hipStream_t>(hipaccStreamOf(ep)), f, static_cast<TO*>(v.data),
                     v.stride, v.offset_x, v.offset_y, v.width, v.height);
  const hipError_t e = hipGetLastError();
  const int st = hipaccCheck(e == hipSuccess ? HIPACC_OK : HIPACC_ERR_HIP, "hipaccLaunchUser");
  if (ep) ep->post_kernel();
  return st;
}

// binned_data(num_bins) of a Kernel<T, TB> class over the image its
// iteration space wrote: a new TB[num_bins] the caller owns.  Blocking.
template <typename F, typename TB, typename T>
TB* hipaccBinUser(const F& f, const HipaccAccessor<T>& is, unsigned num_bins,
                  const HipaccExecutionParameterHip& ep = nullptr) {
  TB* result = new TB[num_bins]();
  const hipacc_view v = hipaccView(is);
  void* bins = nullptr;
  int32_t stride = 0;
  if (hipaccCheck(hipacc_hip_alloc2d(&bins, &stride, (int32_t)num_bins, 1, 4, 0), "hipaccBinUser"))
    return result;
  hipStream_t s = static_cast<hipStream_t>(hipaccStreamOf(ep));
  hipaccCheck(hipacc_hip_fill(bins, 0u, num_bins, 4, s), "hipaccBinUser");
  const long long n = (long long)v.width * v.height;
  const int blocks = (int)std::min<long long>(std::max<long long>((n + 4095) / 4096, 1), 1024);
  if (ep) ep->pre_kernel();
  hipLaunchKernelGGL((hipacc_dev::user_binning_kernel<F, T, TB>), dim3(blocks), dim3(256),
                     (size_t)num_bins * sizeof(TB), s, f, static_cast<const T*>(v.data), v.stride,
                     v.offset_x, v.offset_y, v.width, v.height, static_cast<TB*>(bins), num_bins);
  const hipError_t e = hipGetLastError();
  if (!hipaccCheck(e == hipSuccess ? HIPACC_OK : HIPACC_ERR_HIP, "hipaccBinUser"))
    hipaccCheck(hipacc_hip_download2d(result, bins, (int32_t)num_bins, (int32_t)num_bins, 1, 4, s),
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
  hipLaunchKernelGGL((hipacc_dev::user_reduce_kernel<F, T>), dim3(blocks), dim3(256), 0, s, f,
                     static_cast<const T*>(v.data), v.stride, v.offset_x, v.offset_y, v.width,
                     v.height, static_cast<T*>(part));
  hipLaunchKernelGGL((hipacc_dev::user_reduce_final<F, T>), dim3(1), dim3(64), 0, s, f,
                     static_cast<T*>(part), blocks);
  const hipError_t e = hipGetLastError();
  if (!hipaccCheck(e == hipSuccess ? HIPACC_OK : HIPACC_ERR_HIP, "hipaccReduceUser"))
    hipaccCheck(hipacc_hip_download2d(&result, part, 1, 1, 1, (int32_t)sizeof(T), s),
                "hipaccReduceUser");
  if (ep) ep->post_kernel();
  (void)hipacc_hip_free(part);
  return result;
}

#endif  // HIPACC_HIP_DEVICE_HPP
