// ref_dsl_driver.cpp -- TEST INFRASTRUCTURE ONLY.
//
// A Hipacc DSL program, written against the reference's own DSL headers
// (include path dsl/ of the reference checkout), that runs the reference's
// host emulation of one operator on SplitMix64 pixels and writes the result
// as raw bytes.  It is the "DSL" backend the reference registers as a test
// target (samples-public/CMakeLists.txt.in:36-78), used here to produce the
// golden vectors under tests/golden/ and to check the C restatement in
// hipacc_oracle.c.  Built by oracle/Makefile into oracle/_ref/ref_dsl.
//
// usage: ref_dsl <op> <width> <height> <seed> <boundary 0..4> <out.bin> [args]
//   ops:
//     conv_u8   <k> <sigma>   uchar Gaussian: (uchar)(convolve(SUM, mask()*in(mask)) + 0.5f)
//     conv_f32  <k> <sigma>   float Gaussian: convolve(SUM, mask()*in(mask))
//     sobel_u8  <axis>        short Sobel 3x3: (short)reduce(dom, SUM, mask(dom)*in(dom)) -> int
//     box_u8    <k>           uint box sum: reduce(dom, SUM, in(dom)) -> int
//     min_u8    <k>           uchar erode:  reduce(dom, MIN, in(dom))
//     max_f32   <k>           float dilate: reduce(dom, MAX, in(dom))
//     bilat_f32 <k> <sigma_d> <sigma_r>   the Bilateral_Filter kernel on float pixels
//     red_f32   <0=sum|1=min|2=max>       reduced_data() of a float image
//     red_u8    <1=min|2=max>             reduced_data() of a uchar image
//     hist_u8   <bins>                    binned_data(): bin(pixel*bins/256) = 1, +
//     sub_i8                              point: output() = in1() - in2()  (char)
//     nn_f32    <dst_w> <dst_h>           point over a dst_w x dst_h space reading
//                                         the source through Interpolate::NN
//     lf_f32    <dst_w> <dst_h>           same through Interpolate::LF
//     conv_rgba <k> <sigma>  uchar4 Gaussian (the Gaussian_Blur_RGBA kernel):
//                            convert_uchar4(convolve(SUM, mask()*convert_float4(in(mask))) + 0.5f)
//     conv_f32x4 <k> <sigma> float4 Gaussian: convolve(SUM, mask()*in(mask))
//     box_rgba  <k>          uchar4 box (Box_Blur_RGBA): convert_uchar4(
//                            convert_float4(reduce(dom, SUM, convert_int4(in(dom)))) / (k*k))
//     erode_rgba <k>         uchar4 erode (Erode_RGBA): reduce(dom, MIN, in(dom))
//     sobel_rgba <axis>      int4 Sobel (Sobel_RGBA): reduce(dom, SUM, mask(dom)*convert_int4(in(dom)))
//     interp_f32 / interp_u8 <imode 1..5> <dst_w> <dst_h> <xf> <yf> <cx> <cy> <cw> <ch>
//                                         output() = in(xf, yf) through Interpolate
//                                         NN/LF/B5/CF/L3, the accessor a crop
//                                         (cx, cy, cw, ch; cw = 0: the whole
//                                         image) with the boundary condition
//     fuse_u8 / fuse_f32 <stages>  the kernel chains of the fusion samples
//                            (samples-public/6_Test/Kernel_Fusion_*): each
//                            letter of <stages> one kernel over the previous
//                            one's image, P = point `v = in(); v += 3;
//                            output() = v`, L = local (T)(convolve(5x5
//                            sample mask, SUM, mask()*in(mask)) + 0.5f)
//                            with the boundary condition
#include "hipacc.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "splitmix.h"

using namespace hipacc;
using namespace hipacc::math;

static Boundary g_bmode = Boundary::CLAMP;

template <typename T> static std::vector<T> pixels(int w, int h, uint64_t seed);
template <> std::vector<uchar> pixels<uchar>(int w, int h, uint64_t seed) {
  std::vector<uchar> v((size_t)w * h);
  for (size_t i = 0; i < v.size(); ++i) v[i] = sm64_u8(seed, i);
  return v;
}
template <> std::vector<float> pixels<float>(int w, int h, uint64_t seed) {
  std::vector<float> v((size_t)w * h);
  for (size_t i = 0; i < v.size(); ++i) v[i] = sm64_f32(seed, i);
  return v;
}
template <> std::vector<char> pixels<char>(int w, int h, uint64_t seed) {
  std::vector<char> v((size_t)w * h);
  for (size_t i = 0; i < v.size(); ++i) v[i] = (char)sm64_u8(seed, i);
  return v;
}

static void write_out(const char* path, const void* p, size_t n) {
  FILE* f = std::fopen(path, "wb");
  if (!f || std::fwrite(p, 1, n, f) != n) { std::perror(path); std::exit(2); }
  std::fclose(f);
}

// Normalised Gaussian mask, w(i,j) = exp(-(i^2+j^2)/(2 sigma^2)) / sum, in float.
template <int K> static void gauss(float (&m)[K][K], double sigma) {
  double s = 0.0;
  double t[K][K];
  for (int y = 0; y < K; ++y)
    for (int x = 0; x < K; ++x) {
      const double dx = x - K / 2, dy = y - K / 2;
      t[y][x] = std::exp(-(dx * dx + dy * dy) / (2.0 * sigma * sigma));
      s += t[y][x];
    }
  for (int y = 0; y < K; ++y)
    for (int x = 0; x < K; ++x) m[y][x] = (float)(t[y][x] / s);
}

template <typename T> static BoundaryCondition<T> make_bc(Image<T>& img, MaskBase& m) {
  if (g_bmode == Boundary::CONSTANT) return BoundaryCondition<T>(img, m, g_bmode, T{});
  return BoundaryCondition<T>(img, m, g_bmode);
}

// ---- kernels ---------------------------------------------------------------
class ConvU8 : public Kernel<uchar> {
  Accessor<uchar>& in; Mask<float>& mask;
 public:
  ConvU8(IterationSpace<uchar>& is, Accessor<uchar>& in, Mask<float>& m) : Kernel(is), in(in), mask(m) { add_accessor(&in); }
  void kernel() { output() = (uchar)(convolve(mask, Reduce::SUM, [&]() -> float { return mask() * in(mask); }) + 0.5f); }
};
class ConvF32 : public Kernel<float> {
  Accessor<float>& in; Mask<float>& mask;
 public:
  ConvF32(IterationSpace<float>& is, Accessor<float>& in, Mask<float>& m) : Kernel(is), in(in), mask(m) { add_accessor(&in); }
  void kernel() { output() = convolve(mask, Reduce::SUM, [&]() -> float { return mask() * in(mask); }); }
};
class SobelU8 : public Kernel<short> {
  Accessor<uchar>& in; Domain& dom; Mask<int>& mask;
 public:
  SobelU8(IterationSpace<short>& is, Accessor<uchar>& in, Domain& d, Mask<int>& m) : Kernel(is), in(in), dom(d), mask(m) { add_accessor(&in); }
  void kernel() { output() = (short)(reduce(dom, Reduce::SUM, [&]() -> int { return mask(dom) * in(dom); })); }
};
class BoxU8 : public Kernel<uint> {
  Accessor<uchar>& in; Domain& dom;
 public:
  BoxU8(IterationSpace<uint>& is, Accessor<uchar>& in, Domain& d) : Kernel(is), in(in), dom(d) { add_accessor(&in); }
  void kernel() { output() = reduce(dom, Reduce::SUM, [&]() -> int { return in(dom); }); }
};
class MinU8 : public Kernel<uchar> {
  Accessor<uchar>& in; Domain& dom;
 public:
  MinU8(IterationSpace<uchar>& is, Accessor<uchar>& in, Domain& d) : Kernel(is), in(in), dom(d) { add_accessor(&in); }
  void kernel() { output() = reduce(dom, Reduce::MIN, [&]() -> uchar { return in(dom); }); }
};
class MaxF32 : public Kernel<float> {
  Accessor<float>& in; Domain& dom;
 public:
  MaxF32(IterationSpace<float>& is, Accessor<float>& in, Domain& d) : Kernel(is), in(in), dom(d) { add_accessor(&in); }
  void kernel() { output() = reduce(dom, Reduce::MAX, [&]() -> float { return in(dom); }); }
};
class BilatF32 : public Kernel<float> {
  Accessor<float>& in; Mask<float>& mask; Domain& dom; float sigma_r;
 public:
  BilatF32(IterationSpace<float>& is, Accessor<float>& in, Mask<float>& m, Domain& d, float sr)
      : Kernel(is), in(in), mask(m), dom(d), sigma_r(sr) { add_accessor(&in); }
  void kernel() {
    float c_r = 0.5f / (sigma_r * sigma_r);
    float d = 0.0f, p = 0.0f;
    float center = in();
    iterate(dom, [&]() -> void {
      float diff = in(dom) - center;
      float s = expf(-c_r * diff * diff) * mask(dom);
      d += s;
      p += s * in(dom);
    });
    output() = p / d;
  }
};
class CopyF32 : public Kernel<float> {
  Accessor<float>& in; int op;
 public:
  CopyF32(IterationSpace<float>& is, Accessor<float>& in, int op) : Kernel(is), in(in), op(op) { add_accessor(&in); }
  void kernel() { output() = in(); }
  float reduce(float l, float r) const {
    if (op == 1) return min(l, r);
    if (op == 2) return max(l, r);
    return l + r;
  }
};
class CopyU8 : public Kernel<uchar> {
  Accessor<uchar>& in; int op;
 public:
  CopyU8(IterationSpace<uchar>& is, Accessor<uchar>& in, int op) : Kernel(is), in(in), op(op) { add_accessor(&in); }
  void kernel() { output() = in(); }
  uchar reduce(uchar l, uchar r) const { return op == 1 ? min(l, r) : max(l, r); }
};
class HistU8 : public Kernel<uchar, uint> {
  Accessor<uchar>& in;
 public:
  HistU8(IterationSpace<uchar>& is, Accessor<uchar>& in) : Kernel(is), in(in) { add_accessor(&in); }
  void kernel() { output() = in(); }
  void binning(unsigned int x, unsigned int y, uchar pixel) { bin((unsigned)pixel * num_bins() / 256u) = 1; }
  uint reduce(uint l, uint r) const { return l + r; }
};
class SubI8 : public Kernel<char> {
  Accessor<char>& a; Accessor<char>& b;
 public:
  SubI8(IterationSpace<char>& is, Accessor<char>& a, Accessor<char>& b) : Kernel(is), a(a), b(b) { add_accessor(&a); add_accessor(&b); }
  void kernel() { output() = a() - b(); }
};
class Resample : public Kernel<float> {
  Accessor<float>& in;
 public:
  Resample(IterationSpace<float>& is, Accessor<float>& in) : Kernel(is), in(in) { add_accessor(&in); }
  void kernel() { output() = in(); }
};

// 4-wide vector pixels: the RGBA samples' kernels
class ConvRGBA : public Kernel<uchar4> {
  Accessor<uchar4>& in; Mask<float>& mask;
 public:
  ConvRGBA(IterationSpace<uchar4>& is, Accessor<uchar4>& in, Mask<float>& m) : Kernel(is), in(in), mask(m) { add_accessor(&in); }
  void kernel() {
    float4 sum = convolve(mask, Reduce::SUM, [&]() -> float4 { return mask() * convert_float4(in(mask)); });
    output() = convert_uchar4(sum + 0.5f);
  }
};
class ConvF32x4 : public Kernel<float4> {
  Accessor<float4>& in; Mask<float>& mask;
 public:
  ConvF32x4(IterationSpace<float4>& is, Accessor<float4>& in, Mask<float>& m) : Kernel(is), in(in), mask(m) { add_accessor(&in); }
  void kernel() { output() = convolve(mask, Reduce::SUM, [&]() -> float4 { return mask() * in(mask); }); }
};
class BoxRGBA : public Kernel<uchar4> {
  Accessor<uchar4>& in; Domain& dom; int n;
 public:
  BoxRGBA(IterationSpace<uchar4>& is, Accessor<uchar4>& in, Domain& d, int n) : Kernel(is), in(in), dom(d), n(n) { add_accessor(&in); }
  void kernel() {
    int4 sum = reduce(dom, Reduce::SUM, [&]() -> int4 { return convert_int4(in(dom)); });
    output() = convert_uchar4(convert_float4(sum) / (float)n);
  }
};
class ErodeRGBA : public Kernel<uchar4> {
  Accessor<uchar4>& in; Domain& dom;
 public:
  ErodeRGBA(IterationSpace<uchar4>& is, Accessor<uchar4>& in, Domain& d) : Kernel(is), in(in), dom(d) { add_accessor(&in); }
  void kernel() { output() = reduce(dom, Reduce::MIN, [&]() -> uchar4 { return in(dom); }); }
};
class SobelRGBA : public Kernel<int4> {
  Accessor<uchar4>& in; Domain& dom; Mask<int>& mask;
 public:
  SobelRGBA(IterationSpace<int4>& is, Accessor<uchar4>& in, Domain& d, Mask<int>& m) : Kernel(is), in(in), dom(d), mask(m) { add_accessor(&in); }
  void kernel() { output() = reduce(dom, Reduce::SUM, [&]() -> int4 { return mask(dom) * convert_int4(in(dom)); }); }
};

static std::vector<uchar4> pixels_rgba(int w, int h, uint64_t seed) {
  std::vector<uchar4> v((size_t)w * h);
  for (size_t i = 0; i < v.size(); ++i) {
    v[i].x = sm64_u8(seed, 4 * i); v[i].y = sm64_u8(seed, 4 * i + 1);
    v[i].z = sm64_u8(seed, 4 * i + 2); v[i].w = sm64_u8(seed, 4 * i + 3);
  }
  return v;
}
static std::vector<float4> pixels_f32x4(int w, int h, uint64_t seed) {
  std::vector<float4> v((size_t)w * h);
  for (size_t i = 0; i < v.size(); ++i) {
    v[i].x = sm64_f32(seed, 4 * i); v[i].y = sm64_f32(seed, 4 * i + 1);
    v[i].z = sm64_f32(seed, 4 * i + 2); v[i].w = sm64_f32(seed, 4 * i + 3);
  }
  return v;
}
template <int K> static int run_rgba(const std::string& op, int w, int h, uint64_t seed, double sigma,
                                     const char* path) {
  float coef[K][K];
  gauss<K>(coef, sigma);
  Mask<float> fmask(coef);
  if (op == "conv_f32x4") {
    auto px = pixels_f32x4(w, h, seed);
    Image<float4> in(w, h, px.data()), out(w, h);
    auto bc = make_bc(in, fmask);
    Accessor<float4> acc(bc);
    IterationSpace<float4> is(out);
    ConvF32x4 k(is, acc, fmask);
    k.execute();
    write_out(path, out.data(), (size_t)w * h * 16);
    return 0;
  }
  auto px = pixels_rgba(w, h, seed);
  Image<uchar4> in(w, h, px.data()), out(w, h);
  if (op == "conv_rgba") {
    auto bc = make_bc(in, fmask);
    Accessor<uchar4> acc(bc);
    IterationSpace<uchar4> is(out);
    ConvRGBA k(is, acc, fmask);
    k.execute();
  } else {
    Domain dom(K, K);
    auto bc = make_bc(in, dom);
    Accessor<uchar4> acc(bc);
    IterationSpace<uchar4> is(out);
    if (op == "box_rgba") {
      BoxRGBA k(is, acc, dom, K * K);
      k.execute();
    } else {
      ErodeRGBA k(is, acc, dom);
      k.execute();
    }
  }
  write_out(path, out.data(), (size_t)w * h * 4);
  return 0;
}
static int run_sobel_rgba(int w, int h, uint64_t seed, int axis, const char* path) {
  const int cx[3][3] = {{-1, 0, 1}, {-2, 0, 2}, {-1, 0, 1}};
  const int cy[3][3] = {{-1, -2, -1}, {0, 0, 0}, {1, 2, 1}};
  auto px = pixels_rgba(w, h, seed);
  Image<uchar4> in(w, h, px.data());
  Image<int4> out(w, h);
  Mask<int> mask(axis ? cy : cx);
  Domain dom(mask);
  auto bc = make_bc(in, dom);
  Accessor<uchar4> acc(bc);
  IterationSpace<int4> is(out);
  SobelRGBA k(is, acc, dom, mask);
  k.execute();
  write_out(path, out.data(), (size_t)w * h * 16);
  return 0;
}

template <typename T> class ResampleAt : public Kernel<T> {
  Accessor<T>& in;
  const int xf, yf;
 public:
  ResampleAt(IterationSpace<T>& is, Accessor<T>& in, int xf, int yf)
      : Kernel<T>(is), in(in), xf(xf), yf(yf) { this->add_accessor(&in); }
  void kernel() { this->output() = in(xf, yf); }
};

template <typename T>
static int run_interp(int w, int h, uint64_t seed, int imode, int dw, int dh, int xf, int yf,
                      int cx, int cy, int cw, int ch, const char* path) {
  auto px = pixels<T>(w, h, seed);
  Image<T> in(w, h, px.data()), out(dw, dh);
  // CONSTANT needs a constant (the DSL reads T{} whatever value is given)
  BoundaryCondition<T> bc = g_bmode == Boundary::CONSTANT
                                ? BoundaryCondition<T>(in, 1, 1, g_bmode, T(0))
                                : BoundaryCondition<T>(in, 1, 1, g_bmode);
  if (cw <= 0) { cx = 0; cy = 0; cw = w; ch = h; }
  Accessor<T> acc(bc, cw, ch, cx, cy, (Interpolate)imode);
  IterationSpace<T> is(out);
  ResampleAt<T> k(is, acc, xf, yf);
  k.execute();
  write_out(path, out.data(), (size_t)dw * dh * sizeof(T));
  return 0;
}

// ---- fusion chains -----------------------------------------------------------
// The point and local kernels of samples-public/6_Test/Kernel_Fusion_P2L/src/
// main.cpp:43-78 (the same in the P2P, L2P, L2L and Mixed samples), and
// their 5x5 mask (main.cpp:91-97).
template <typename T> class AddThree : public Kernel<T> {
  Accessor<T>& in;
 public:
  AddThree(IterationSpace<T>& is, Accessor<T>& in) : Kernel<T>(is), in(in) { this->add_accessor(&in); }
  void kernel() {
    T v = in();
    v += 3;
    this->output() = v;
  }
};
template <typename T> class GaussRound : public Kernel<T> {
  Accessor<T>& in;
  Mask<float>& mask;
 public:
  GaussRound(IterationSpace<T>& is, Accessor<T>& in, Mask<float>& m)
      : Kernel<T>(is), in(in), mask(m) { this->add_accessor(&in); }
  void kernel() {
    this->output() = (T)(this->convolve(mask, Reduce::SUM, [&]() -> float { return mask() * in(mask); }) + 0.5f);
  }
};
static const float kFuseMask[5][5] = {
    {0.005008f, 0.017300f, 0.026151f, 0.017300f, 0.005008f},
    {0.017300f, 0.059761f, 0.090339f, 0.059761f, 0.017300f},
    {0.026151f, 0.090339f, 0.136565f, 0.090339f, 0.026151f},
    {0.017300f, 0.059761f, 0.090339f, 0.059761f, 0.017300f},
    {0.005008f, 0.017300f, 0.026151f, 0.017300f, 0.005008f}};

template <typename T>
static int run_fuse(int w, int h, uint64_t seed, const std::string& stages, const char* path) {
  std::vector<T> cur = pixels<T>(w, h, seed);
  Mask<float> mask(kFuseMask);
  for (char c : stages) {
    Image<T> in(w, h, cur.data()), out(w, h);
    IterationSpace<T> is(out);
    if (c == 'P') {
      Accessor<T> acc(in);
      AddThree<T> k(is, acc);
      k.execute();
    } else if (c == 'L') {
      auto bc = make_bc(in, mask);
      Accessor<T> acc(bc);
      GaussRound<T> k(is, acc, mask);
      k.execute();
    } else {
      std::fprintf(stderr, "fuse: stage %c\n", c);
      return 2;
    }
    const T* o = out.data();
    cur.assign(o, o + (size_t)w * h);
  }
  write_out(path, cur.data(), (size_t)w * h * sizeof(T));
  return 0;
}

// ---- drivers ----------------------------------------------------------------
template <int K> static int run_conv_u8(int w, int h, uint64_t seed, double sigma, const char* path) {
  float coef[K][K];
  gauss<K>(coef, sigma);
  auto px = pixels<uchar>(w, h, seed);
  Image<uchar> in(w, h, px.data()), out(w, h);
  Mask<float> mask(coef);
  auto bc = make_bc(in, mask);
  Accessor<uchar> acc(bc);
  IterationSpace<uchar> is(out);
  ConvU8 k(is, acc, mask);
  k.execute();
  write_out(path, out.data(), (size_t)w * h);
  return 0;
}
template <int K> static int run_conv_f32(int w, int h, uint64_t seed, double sigma, const char* path) {
  float coef[K][K];
  gauss<K>(coef, sigma);
  auto px = pixels<float>(w, h, seed);
  Image<float> in(w, h, px.data()), out(w, h);
  Mask<float> mask(coef);
  auto bc = make_bc(in, mask);
  Accessor<float> acc(bc);
  IterationSpace<float> is(out);
  ConvF32 k(is, acc, mask);
  k.execute();
  write_out(path, out.data(), (size_t)w * h * 4);
  return 0;
}
static int run_sobel_u8(int w, int h, uint64_t seed, int axis, const char* path) {
  const int cx[3][3] = {{-1, 0, 1}, {-2, 0, 2}, {-1, 0, 1}};
  const int cy[3][3] = {{-1, -2, -1}, {0, 0, 0}, {1, 2, 1}};
  auto px = pixels<uchar>(w, h, seed);
  Image<uchar> in(w, h, px.data());
  Image<short> out(w, h);
  Mask<int> mask(axis ? cy : cx);
  Domain dom(mask);
  auto bc = make_bc(in, dom);
  Accessor<uchar> acc(bc);
  IterationSpace<short> is(out);
  SobelU8 k(is, acc, dom, mask);
  k.execute();
  write_out(path, out.data(), (size_t)w * h * 2);
  return 0;
}
template <int K> static int run_box_u8(int w, int h, uint64_t seed, const char* path) {
  auto px = pixels<uchar>(w, h, seed);
  Image<uchar> in(w, h, px.data());
  Image<uint> out(w, h);
  Domain dom(K, K);
  auto bc = make_bc(in, dom);
  Accessor<uchar> acc(bc);
  IterationSpace<uint> is(out);
  BoxU8 k(is, acc, dom);
  k.execute();
  write_out(path, out.data(), (size_t)w * h * 4);
  return 0;
}
template <int K> static int run_min_u8(int w, int h, uint64_t seed, const char* path) {
  auto px = pixels<uchar>(w, h, seed);
  Image<uchar> in(w, h, px.data()), out(w, h);
  Domain dom(K, K);
  auto bc = make_bc(in, dom);
  Accessor<uchar> acc(bc);
  IterationSpace<uchar> is(out);
  MinU8 k(is, acc, dom);
  k.execute();
  write_out(path, out.data(), (size_t)w * h);
  return 0;
}
template <int K> static int run_max_f32(int w, int h, uint64_t seed, const char* path) {
  auto px = pixels<float>(w, h, seed);
  Image<float> in(w, h, px.data()), out(w, h);
  Domain dom(K, K);
  auto bc = make_bc(in, dom);
  Accessor<float> acc(bc);
  IterationSpace<float> is(out);
  MaxF32 k(is, acc, dom);
  k.execute();
  write_out(path, out.data(), (size_t)w * h * 4);
  return 0;
}
template <int K> static int run_bilat(int w, int h, uint64_t seed, double sd, float sr, const char* path) {
  float coef[K][K];
  for (int y = 0; y < K; ++y)
    for (int x = 0; x < K; ++x) {
      const double dx = x - K / 2, dy = y - K / 2;
      coef[y][x] = (float)std::exp(-(dx * dx + dy * dy) / (2.0 * sd * sd));
    }
  auto px = pixels<float>(w, h, seed);
  Image<float> in(w, h, px.data()), out(w, h);
  Mask<float> mask(coef);
  Domain dom(mask);
  auto bc = make_bc(in, dom);
  Accessor<float> acc(bc);
  IterationSpace<float> is(out);
  BilatF32 k(is, acc, mask, dom, sr);
  k.execute();
  write_out(path, out.data(), (size_t)w * h * 4);
  return 0;
}

int main(int argc, char** argv) {
  if (argc < 7) {
    std::fprintf(stderr, "usage: %s op w h seed boundary out.bin [args]\n", argv[0]);
    return 2;
  }
  const std::string op = argv[1];
  const int w = std::atoi(argv[2]), h = std::atoi(argv[3]);
  const uint64_t seed = std::strtoull(argv[4], nullptr, 10);
  g_bmode = (Boundary)std::atoi(argv[5]);
  const char* path = argv[6];
  auto arg = [&](int i) { return i + 7 < argc ? argv[i + 7] : "0"; };
  if (op == "conv_u8" || op == "conv_f32") {
    const int k = std::atoi(arg(0));
    const double s = std::atof(arg(1));
    const bool u8 = op == "conv_u8";
    switch (k) {
      case 3: return u8 ? run_conv_u8<3>(w, h, seed, s, path) : run_conv_f32<3>(w, h, seed, s, path);
      case 5: return u8 ? run_conv_u8<5>(w, h, seed, s, path) : run_conv_f32<5>(w, h, seed, s, path);
      case 7: return u8 ? run_conv_u8<7>(w, h, seed, s, path) : run_conv_f32<7>(w, h, seed, s, path);
      case 9: return u8 ? run_conv_u8<9>(w, h, seed, s, path) : run_conv_f32<9>(w, h, seed, s, path);
      case 13: return u8 ? run_conv_u8<13>(w, h, seed, s, path) : run_conv_f32<13>(w, h, seed, s, path);
    }
  } else if (op == "sobel_u8") {
    return run_sobel_u8(w, h, seed, std::atoi(arg(0)), path);
  } else if (op == "box_u8") {
    switch (std::atoi(arg(0))) {
      case 3: return run_box_u8<3>(w, h, seed, path);
      case 5: return run_box_u8<5>(w, h, seed, path);
      case 7: return run_box_u8<7>(w, h, seed, path);
    }
  } else if (op == "min_u8") {
    switch (std::atoi(arg(0))) {
      case 3: return run_min_u8<3>(w, h, seed, path);
      case 5: return run_min_u8<5>(w, h, seed, path);
    }
  } else if (op == "max_f32") {
    switch (std::atoi(arg(0))) {
      case 3: return run_max_f32<3>(w, h, seed, path);
      case 5: return run_max_f32<5>(w, h, seed, path);
    }
  } else if (op == "bilat_f32") {
    const double sd = std::atof(arg(1));
    const float sr = (float)std::atof(arg(2));
    switch (std::atoi(arg(0))) {
      case 3: return run_bilat<3>(w, h, seed, sd, sr, path);
      case 5: return run_bilat<5>(w, h, seed, sd, sr, path);
      case 7: return run_bilat<7>(w, h, seed, sd, sr, path);
      case 13: return run_bilat<13>(w, h, seed, sd, sr, path);
    }
  } else if (op == "red_f32") {
    auto px = pixels<float>(w, h, seed);
    Image<float> in(w, h, px.data()), out(w, h);
    Accessor<float> acc(in);
    IterationSpace<float> is(out);
    CopyF32 k(is, acc, std::atoi(arg(0)));
    k.execute();
    const float r = k.reduced_data();
    write_out(path, &r, 4);
    return 0;
  } else if (op == "red_u8") {
    auto px = pixels<uchar>(w, h, seed);
    Image<uchar> in(w, h, px.data()), out(w, h);
    Accessor<uchar> acc(in);
    IterationSpace<uchar> is(out);
    CopyU8 k(is, acc, std::atoi(arg(0)));
    k.execute();
    const uchar r = k.reduced_data();
    write_out(path, &r, 1);
    return 0;
  } else if (op == "hist_u8") {
    const unsigned bins = (unsigned)std::atoi(arg(0));
    auto px = pixels<uchar>(w, h, seed);
    Image<uchar> in(w, h, px.data()), out(w, h);
    Accessor<uchar> acc(in);
    IterationSpace<uchar> is(out);
    HistU8 k(is, acc);
    k.execute();
    uint* r = k.binned_data(bins);
    write_out(path, r, bins * 4);
    delete[] r;
    return 0;
  } else if (op == "sub_i8") {
    auto pa = pixels<char>(w, h, seed);
    auto pb = pixels<char>(w, h, seed + 1);
    Image<char> a(w, h, pa.data()), b(w, h, pb.data()), out(w, h);
    Accessor<char> aa(a), ab(b);
    IterationSpace<char> is(out);
    SubI8 k(is, aa, ab);
    k.execute();
    write_out(path, out.data(), (size_t)w * h);
    return 0;
  } else if (op == "nn_f32" || op == "lf_f32") {
    const int dw = std::atoi(arg(0)), dh = std::atoi(arg(1));
    auto px = pixels<float>(w, h, seed);
    Image<float> in(w, h, px.data()), out(dw, dh);
    Accessor<float> acc(in, op == "nn_f32" ? Interpolate::NN : Interpolate::LF);
    IterationSpace<float> is(out);
    Resample k(is, acc);
    k.execute();
    write_out(path, out.data(), (size_t)dw * dh * 4);
    return 0;
  }
  if (op == "conv_rgba" || op == "conv_f32x4" || op == "box_rgba" || op == "erode_rgba") {
    const int k = std::atoi(arg(0));
    const double s = std::atof(arg(1));
    switch (k) {
      case 3: return run_rgba<3>(op, w, h, seed, s, path);
      case 5: return run_rgba<5>(op, w, h, seed, s, path);
      case 7: return run_rgba<7>(op, w, h, seed, s, path);
    }
  }
  if (op == "sobel_rgba") return run_sobel_rgba(w, h, seed, std::atoi(arg(0)), path);
  if (op == "interp_f32" || op == "interp_u8") {
    int v[9];
    for (int i = 0; i < 9; ++i) v[i] = std::atoi(arg(i));
    if (op == "interp_f32")
      return run_interp<float>(w, h, seed, v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7], v[8], path);
    return run_interp<uchar>(w, h, seed, v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7], v[8], path);
  }
  if (op == "fuse_u8") return run_fuse<uchar>(w, h, seed, arg(0), path);
  if (op == "fuse_f32") return run_fuse<float>(w, h, seed, arg(0), path);
  std::fprintf(stderr, "unknown op/size: %s\n", op.c_str());
  return 2;
}
