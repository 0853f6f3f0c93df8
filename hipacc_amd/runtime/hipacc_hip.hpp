// hipacc_hip.hpp -- C++ runtime for Hipacc host code on MI355X (gfx950).
//
// The functions and classes generated host code uses, with the names of the
// reference's CUDA runtime (runtime/hipacc_cu.hpp:86-372,
// runtime/hipacc_base.hpp:54-191): images with a host mirror
// (hipaccCreateMemory / hipaccWriteMemory / hipaccReadMemory), accessors,
// execution parameters carrying a stream, pyramids and their traversal,
// hipacc_last_kernel_timing().  Instead of launching one generated kernel per
// DSL operator, each operator is a call of libhipacc_gfx950.so
// (include/hipacc_hip.h): hipaccLaunchLocal for convolve()/reduce(Domain)
// /iterate() kernels, hipaccLaunchSeparable for a pair of 1D convolutions,
// hipaccLaunchPoint for point operators, hipaccLaunchBilateral, and
// hipaccApplyReduction / hipaccApplyBinning (hipacc_hip_red.hpp) for global
// operators.
//
// Needs no HIP headers: streams are passed as void* (a hipStream_t).  Link
// with -lhipacc_gfx950.  Errors are logged in the reference's style
// (hipaccRuntimeLogTrivial) and execution continues, as the reference does
// (runtime/hipacc_base.hpp:110-129).
#ifndef HIPACC_HIP_HPP
#define HIPACC_HIP_HPP

#include <algorithm>
#include <cassert>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <iostream>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "hipacc_hip.h"

// ---------------------------------------------------------------------------
// Logging and status checks
// ---------------------------------------------------------------------------
enum class hipaccRuntimeLogLevel { INFO, WARNING, ERROR };

inline int hipaccRuntimeLogTrivial(hipaccRuntimeLogLevel level, const std::string& message) {
  if (level == hipaccRuntimeLogLevel::INFO)
    std::cout << message << std::endl;
  else
    std::cerr << message << std::endl;
  return 0;
}

// Logs a failed library call; returns the status.
inline int hipaccCheck(int status, const std::string& what) {
  if (status != HIPACC_OK)
    hipaccRuntimeLogTrivial(hipaccRuntimeLogLevel::ERROR,
                            "ERROR: " + what + " (" + std::to_string(status) + "): " +
                                hipacc_hip_strerror(status) + ": " + hipacc_hip_last_error());
  return status;
}

// ---------------------------------------------------------------------------
// Timing (hipacc_last_kernel_timing, runtime/hipacc_base.hpp:54-66)
// ---------------------------------------------------------------------------
inline float hipacc_last_kernel_timing() { return hipacc_hip_last_kernel_timing(); }
// Kernel timing synchronises after every launch (the reference's
// -time-kernels); off by default.
inline void hipaccSetTiming(bool on) { hipacc_hip_set_timing(on ? 1 : 0); }

// ---------------------------------------------------------------------------
// Pixel types, boundary and reduction modes
// ---------------------------------------------------------------------------
template <typename T> struct hipacc_pixel_code;
template <> struct hipacc_pixel_code<unsigned char> { static constexpr int value = HIPACC_U8; };
template <> struct hipacc_pixel_code<signed char> { static constexpr int value = HIPACC_I8; };
template <> struct hipacc_pixel_code<char> { static constexpr int value = HIPACC_I8; };
template <> struct hipacc_pixel_code<unsigned short> { static constexpr int value = HIPACC_U16; };
template <> struct hipacc_pixel_code<short> { static constexpr int value = HIPACC_I16; };
template <> struct hipacc_pixel_code<unsigned int> { static constexpr int value = HIPACC_U32; };
template <> struct hipacc_pixel_code<int> { static constexpr int value = HIPACC_I32; };
template <> struct hipacc_pixel_code<float> { static constexpr int value = HIPACC_F32; };

// Same values as hipacc::Boundary / hipacc::Reduce (dsl/image.hpp:46-52,
// dsl/kernel.hpp:48-54).
enum class HipaccBoundary : int { UNDEFINED = 0, CLAMP, REPEAT, MIRROR, CONSTANT };
enum class HipaccReduce : int { SUM = 0, MIN, MAX, PROD };

// ---------------------------------------------------------------------------
// Images (runtime/hipacc_cu.hpp:91-160)
// ---------------------------------------------------------------------------
template <typename T>
class HipaccImageHipBase {
 public:
  using pixel_type = T;
  virtual ~HipaccImageHipBase() = default;
  virtual const T* get_device_memory() const = 0;
  virtual T* get_device_memory() = 0;
  virtual const T* get_host_memory() const = 0;
  virtual T* get_host_memory() = 0;
  virtual int get_width() const = 0;
  virtual int get_height() const = 0;
  virtual int get_stride() const = 0;
  virtual int get_alignment() const = 0;
  virtual int get_pixel_size() const = 0;
};

template <typename T>
using HipaccImageHip = std::shared_ptr<HipaccImageHipBase<T>>;

template <typename T>
class HipaccImageHipRaw final : public HipaccImageHipBase<T> {
  const int width_, height_, stride_, alignment_;
  T* const device_mem_;
  std::unique_ptr<T[]> host_mem_;

 public:
  HipaccImageHipRaw(int width, int height, int stride, int alignment, T* mem)
      : width_(width), height_(height), stride_(stride), alignment_(alignment), device_mem_(mem),
        host_mem_(new T[(size_t)width * height]) {
    std::fill(host_mem_.get(), host_mem_.get() + (size_t)width * height, T{});
  }
  ~HipaccImageHipRaw() override { hipaccCheck(hipacc_hip_free(device_mem_), "hipacc_hip_free"); }
  const T* get_device_memory() const override { return device_mem_; }
  T* get_device_memory() override { return device_mem_; }
  const T* get_host_memory() const override { return host_mem_.get(); }
  T* get_host_memory() override { return host_mem_.get(); }
  int get_width() const override { return width_; }
  int get_height() const override { return height_; }
  int get_stride() const override { return stride_; }
  int get_alignment() const override { return alignment_; }
  int get_pixel_size() const override { return (int)sizeof(T); }
};

// ---------------------------------------------------------------------------
// Accessors and iteration spaces (runtime/hipacc_base.hpp:72-82,
// runtime/hipacc_cu.hpp:162-186): a window of an image.
// ---------------------------------------------------------------------------
struct HipaccAccessorBase {
  int width, height, offset_x, offset_y;
  HipaccAccessorBase(int w, int h, int ox = 0, int oy = 0)
      : width(w), height(h), offset_x(ox), offset_y(oy) {}
};

template <typename T>
struct HipaccAccessor : public HipaccAccessorBase {
  HipaccImageHip<T> img;
  HipaccAccessor(const HipaccImageHip<T>& i, int w, int h, int ox = 0, int oy = 0)
      : HipaccAccessorBase(w, h, ox, oy), img(i) {}
  explicit HipaccAccessor(const HipaccImageHip<T>& i)
      : HipaccAccessorBase(i->get_width(), i->get_height(), 0, 0), img(i) {}
};

template <typename T>
inline HipaccAccessor<T> hipaccMakeAccessor(const HipaccImageHip<T>& img) {
  return HipaccAccessor<T>(img);
}
template <typename T>
inline HipaccAccessor<T> hipaccMakeAccessor(const HipaccImageHip<T>& img, int width, int height,
                                            int offset_x = 0, int offset_y = 0) {
  return HipaccAccessor<T>(img, width, height, offset_x, offset_y);
}

// The C-ABI view of an accessor (dtype may be overridden for same-size
// reinterpretation).
template <typename T>
inline hipacc_view hipaccView(const HipaccAccessor<T>& acc) {
  hipacc_view v;
  std::memset(&v, 0, sizeof(v));
  v.data = const_cast<T*>(acc.img->get_device_memory());
  v.dtype = hipacc_pixel_code<T>::value;
  v.stride = acc.img->get_stride();
  v.img_width = acc.img->get_width();
  v.img_height = acc.img->get_height();
  v.width = acc.width;
  v.height = acc.height;
  v.offset_x = acc.offset_x;
  v.offset_y = acc.offset_y;
  v.batch = 1;
  return v;
}

// ---------------------------------------------------------------------------
// Execution parameters (runtime/hipacc_cu.hpp:234-245): the stream a kernel
// runs on, plus hooks around it.
// ---------------------------------------------------------------------------
class HipaccExecutionParameterHipBase {
  void* stream_ = nullptr;

 protected:
  void set_stream(void* s) { stream_ = s; }

 public:
  virtual ~HipaccExecutionParameterHipBase() = default;
  void* get_stream() const { return stream_; }
  virtual void pre_kernel() {}
  virtual void post_kernel() {}
};
using HipaccExecutionParameterHip = std::shared_ptr<HipaccExecutionParameterHipBase>;

// An execution parameter for a given stream.
class HipaccStreamParameter final : public HipaccExecutionParameterHipBase {
 public:
  explicit HipaccStreamParameter(void* stream) { set_stream(stream); }
};

inline void* hipaccStreamOf(const HipaccExecutionParameterHip& ep) {
  return ep ? ep->get_stream() : nullptr;
}

// ---------------------------------------------------------------------------
// Device initialisation (hipaccInitCUDA, runtime/hipacc_cu_standalone.hpp)
// ---------------------------------------------------------------------------
inline void hipaccInitHIP(int device = 0) {
  hipaccCheck(hipacc_hip_init(device), "hipacc_hip_init");
}

// ---------------------------------------------------------------------------
// Masks: the coefficient table of a Mask / the holes of a Domain (replaces
// hipaccWriteSymbol / hipaccWriteDomainFromMask, runtime/hipacc_cu.tpp:
// 278-309; the tables travel in the kernel arguments).
// ---------------------------------------------------------------------------
template <typename M>
struct HipaccMask {
  int size_x, size_y;
  std::vector<M> coef;         // row-major size_y x size_x
  std::vector<uint8_t> domain; // nonzero: tap active
  HipaccMask(int sx, int sy) : size_x(sx), size_y(sy), coef((size_t)sx * sy, M(1)),
                               domain((size_t)sx * sy, 1) {}
  HipaccMask(int sx, int sy, const M* host) : HipaccMask(sx, sy) { write(host); }
  void write(const M* host) { std::copy(host, host + (size_t)size_x * size_y, coef.begin()); }
  // Domain(mask): holes where the coefficient is 0 (dsl/mask.hpp:237-250)
  void domain_from_coef() {
    for (size_t i = 0; i < coef.size(); ++i) domain[i] = coef[i] != M(0);
  }
  // Domain(array): enabled where the row-major size_y x size_x array is
  // nonzero (dsl/mask.hpp:155-164).
  template <typename E> void domain_write(const E* host) {
    for (size_t i = 0; i < domain.size(); ++i) domain[i] = host[i] != E(0);
  }
  // Element (xf, yf), offsets from the centre, of the domain and of the
  // coefficients: the DSL's dom(xf, yf) = v and mask(xf, yf) = v, at row
  // size_y / 2 + yf and column size_x / 2 + xf as the reference compiler
  // places a domain assignment (lib/Rewrite/Rewrite.cpp:1650-1678).
  uint8_t& domain_at(int xf, int yf) {
    return domain.at((size_t)(yf + size_y / 2) * size_x + xf + size_x / 2);
  }
  M& coef_at(int xf, int yf) {
    return coef.at((size_t)(yf + size_y / 2) * size_x + xf + size_x / 2);
  }
};

template <typename M>
inline void hipaccWriteSymbol(HipaccMask<M>& mask, const M* host) { mask.write(host); }
template <typename M>
inline void hipaccWriteDomainFromMask(HipaccMask<M>& mask, const M* host) {
  mask.write(host);
  mask.domain_from_coef();
}

// The mask with its coefficients converted to To (the accumulator type of a
// kernel's lambda: float or int) and the same holes.
template <typename To, typename From>
inline HipaccMask<To> hipaccMaskCast(const HipaccMask<From>& m) {
  HipaccMask<To> r(m.size_x, m.size_y);
  for (size_t i = 0; i < m.coef.size(); ++i) r.coef[i] = (To)m.coef[i];
  r.domain = m.domain;
  return r;
}

// Options of a local operator's launch.
struct HipaccLocalOptions {
  HipaccReduce reduce = HipaccReduce::SUM;
  HipaccBoundary boundary = HipaccBoundary::CLAMP;
  double const_val = 0.0;
  bool use_weights = true;     // false: tap = pixel (reduce(Domain, ...))
  bool round = false;          // (T)(acc + 0.5f) instead of (T)acc
  float post_div = 0.0f;       // (T)((float)acc / post_div)
  bool fp_contract = true;     // acc = fma(w, x, acc)
};

#include "hipacc_hip.tpp"
#include "hipacc_hip_red.hpp"

#endif  // HIPACC_HIP_HPP
