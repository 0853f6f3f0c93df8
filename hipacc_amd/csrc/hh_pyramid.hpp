// hh_pyramid.hpp -- argument blocks of the resampling point operators and
// of the batched pyramid level kernel.
#pragma once

#include "hh_local.hpp"

namespace hh {

// Interpolation::interpolate's coordinate map (dsl/image.hpp:390-407), in
// float without contraction: s = in_extent / (float)out_extent,
// m = (off + s/2) + s*x, with off the accessor's offset.
HH_HD float interp_map(float s, int x, int off = 0) {
  const float base = (float)off + s / 2.0f;
  const float sx = s * (float)x;
  return base + sx;
}
// LF's integer and fractional parts of the mapped position (the position
// moved from cell-centred to absolute, clamped at 0).
HH_HD void interp_lf(float m, int& i, float& f) {
  float b = m - 0.5f;
  if (b < 0.0f) b = 0.0f;
  i = (int)b;
  f = b - (float)i;
}
// Bilinear combination in the DSL's evaluation order.
HH_HD float interp_bilinear(float fx, float fy, float p00, float p10, float p01, float p11) {
  const float gx = 1.0f - fx, gy = 1.0f - fy;
  const float w00 = gx * gy, w10 = fx * gy, w01 = gx * fy, w11 = fx * fy;
  const float t00 = w00 * p00, t10 = w10 * p10, t01 = w01 * p01, t11 = w11 * p11;
  const float s0 = t00 + t10;
  const float s1 = s0 + t01;
  return s1 + t11;
}

// Interpolation weights of the reference's Interpolation class
// (dsl/image.hpp:316-382), in its operation order: binomial_5, Keys' cubic
// convolution with a = -0.5 (bicubic_spline), and Lanczos-3 evaluated in
// double and rounded once to float.
HH_HD float interp_w_b5(float d) {
  d = fabsf(d);
  if (d < 0.5f) return 6.0f / 8.0f;
  if (d < 1.0f) return 4.0f / 8.0f;
  if (d < 1.5f) return 1.0f / 8.0f;
  return 0.0f;
}
HH_HD float interp_w_cf(float d) {
  const float a = -0.5f;
  d = fabsf(d);
  if (d < 1.0f) {
    const float t3 = (a + 2.0f) * d * d * d;
    const float t2 = (a + 3.0f) * d * d;
    return (t3 - t2) + 1.0f;
  }
  if (d < 2.0f) {
    const float t3 = a * d * d * d;
    const float t2 = 5.0f * a * d * d;
    const float t1 = 8.0f * a * d;
    const float t0 = 4.0f * a;
    return ((t3 - t2) + t1) - t0;
  }
  return 0.0f;
}
HH_HD float interp_w_l3(float d) {
  const float l = 3.0f;
  d = fabsf(d);
  if (d == 0.0f) return 1.0f;
  if (d < l) {
    const double pi = 3.141592653589793;   // atan(1.0) * 4, as the reference
    const double num = (double)l * (sin(pi * d / l) * sin(pi * d));
    const double den = pi * pi * d * d;
    return (float)(num / den);
  }
  return 0.0f;
}

// The interpolated value at mapped position (mx, my) (dsl/image.hpp:
// 390-532); read(x, y) returns the pixel at absolute image position (x, y)
// through the accessor's boundary handling, as a float.  NN reads the pixel
// at the truncated position; LF is the bilinear combination; B5 / CF / L3
// sum rows of NX taps (left to right, one rounding per product and add)
// weighted per row:
//   B5: frac + 0.5, taps x_int + k, weight b5(frac - k);
//   CF: taps x_int - 1 + k, weight cf((frac - 1) + k);
//   L3: taps x_int - 2 + k, rows y_int - 1 + k, weight l3((frac - 2) + k),
//       and in the second row the fifth tap takes the sixth tap's weight.
template <int MODE, typename R>
HH_HD float interp_sample(const R& read, float mx, float my) {
  if constexpr (MODE == HIPACC_INTERP_NN) {
    return read((int)mx, (int)my);
  } else if constexpr (MODE == HIPACC_INTERP_LF) {
    int xi, yi;
    float fx, fy;
    interp_lf(mx, xi, fx);
    interp_lf(my, yi, fy);
    return interp_bilinear(fx, fy, read(xi, yi), read(xi + 1, yi), read(xi, yi + 1),
                           read(xi + 1, yi + 1));
  } else {
    int xi, yi;
    float fx, fy;
    interp_lf(mx, xi, fx);
    interp_lf(my, yi, fy);
    constexpr int NX = MODE == HIPACC_INTERP_L3 ? 6 : 4;
    constexpr int NY = MODE == HIPACC_INTERP_L3 ? 6 : 4;
    float bx, by;
    int x0, y0;
    if constexpr (MODE == HIPACC_INTERP_B5) {
      bx = fx + 0.5f;
      by = fy + 0.5f;
      x0 = xi;
      y0 = yi;
    } else if constexpr (MODE == HIPACC_INTERP_CF) {
      bx = fx - 1.0f;
      by = fy - 1.0f;
      x0 = xi - 1;
      y0 = yi - 1;
    } else {
      bx = fx - 2.0f;
      by = fy - 2.0f;
      x0 = xi - 2;
      y0 = yi - 1;
    }
    float v = 0.0f;
#pragma unroll
    for (int r = 0; r < NY; ++r) {
      float row = 0.0f;
#pragma unroll
      for (int k = 0; k < NX; ++k) {
        float w;
        if constexpr (MODE == HIPACC_INTERP_B5) w = interp_w_b5(bx - (float)k);
        else if constexpr (MODE == HIPACC_INTERP_CF) w = interp_w_cf(bx + (float)k);
        else w = interp_w_l3(bx + (float)((r == 1 && k == 4) ? 5 : k));
        const float t = read(x0 + k, y0 + r) * w;
        row = k == 0 ? t : row + t;
      }
      float wy;
      if constexpr (MODE == HIPACC_INTERP_B5) wy = interp_w_b5(by - (float)r);
      else if constexpr (MODE == HIPACC_INTERP_CF) wy = interp_w_cf(by + (float)r);
      else wy = interp_w_l3(by + (float)r);
      const float t = row * wy;
      v = r == 0 ? t : v + t;
    }
    return v;
  }
}

struct ResampleArgs {
  const void* in;              // in_dtype pixels (F32 or U8)
  const float* in2;
  void* out;                   // out_dtype pixels (the input's type)
  int in_dtype, out_dtype;
  int xf, yf;                  // accessor call offsets in(xf, yf)
  int in_stride, in2_stride, out_stride;
  int in_imgw, in_imgh;
  int lo_x, lo_y, hi_x, hi_y;      // accessor window
  int in2_ox, in2_oy, out_ox, out_oy;
  int w, h;                        // iteration space
  float sx, sy;                    // window extent / iteration-space extent
  int interp, op, boundary;
  float cval;
};
int launch_resample(const ResampleArgs& a, hipStream_t s);

// One image of a batched pyramid, as the level kernel sees it.
constexpr int kMaxLevels = HIPACC_MAX_LEVELS;
struct PyrImage {
  int w[kMaxLevels], h[kMaxLevels];
  float* g[kMaxLevels];
  int gs[kMaxLevels];
  float* lap[kMaxLevels];
  int ls[kMaxLevels];
};
struct PyrTile {
  int img, x0, y0;
};
// The regions one tile of level l stages.  The tile covers L[l] pixels
// [x0, x0+TW) x [y0, y0+TH) of a W x H level whose next level is W1 x H1.
// u*: the G[l+1] points it needs -- the LF taps of its pixels plus the G[l+1]
// block it owns (o*: columns x0/2 .. , rows y0/2 ..); r*: the G[l] region the
// 5x5 blur taps of those points read (NN point of G[l+1] pixel u is
// (int)interp_map(W / W1, u)).  The same float arithmetic runs on the host,
// where the plan proves that every tile fits its kernel's LDS capacity.
struct PyrRegion {
  int ux0, uy0, uw, uh;
  int ox0, ox1, oy0, oy1;
  int rx0, ry0, rw, rh;
};
HH_HD PyrRegion pyr_tile_region(int W, int H, int W1, int H1, int x0, int y0, int TW, int TH) {
  PyrRegion r;
  const int xe = (x0 + TW < W ? x0 + TW : W) - 1, ye = (y0 + TH < H ? y0 + TH : H) - 1;
  const float su = (float)W1 / (float)W, sv = (float)H1 / (float)H;
  int ux0, ux1, uy0, uy1;
  float f;
  interp_lf(interp_map(su, x0), ux0, f);
  interp_lf(interp_map(su, xe), ux1, f);
  interp_lf(interp_map(sv, y0), uy0, f);
  interp_lf(interp_map(sv, ye), uy1, f);
  ux1 = ux1 + 1 < W1 - 1 ? ux1 + 1 : W1 - 1;
  uy1 = uy1 + 1 < H1 - 1 ? uy1 + 1 : H1 - 1;
  r.ox0 = x0 / 2;
  r.ox1 = (x0 / 2 + TW / 2 < W1 ? x0 / 2 + TW / 2 : W1) - 1;
  r.oy0 = y0 / 2;
  r.oy1 = (y0 / 2 + TH / 2 < H1 ? y0 / 2 + TH / 2 : H1) - 1;
  ux0 = ux0 < r.ox0 ? ux0 : r.ox0;
  uy0 = uy0 < r.oy0 ? uy0 : r.oy0;
  ux1 = ux1 > r.ox1 ? ux1 : r.ox1;
  uy1 = uy1 > r.oy1 ? uy1 : r.oy1;
  r.ux0 = ux0;
  r.uy0 = uy0;
  r.uw = ux1 - ux0 + 1;
  r.uh = uy1 - uy0 + 1;
  const float sd = (float)W / (float)W1, td = (float)H / (float)H1;
  r.rx0 = (int)interp_map(sd, ux0) - 2;
  r.ry0 = (int)interp_map(td, uy0) - 2;
  r.rw = (int)interp_map(sd, ux1) + 2 - r.rx0 + 1;
  r.rh = (int)interp_map(td, uy1) + 2 - r.ry0 + 1;
  return r;
}

// LDS capacity of a TW x TH tile kernel whose images have NN step
// s = W / W1 <= SNUM / 4: the G[l+1] region (UW x UH) and the G[l] region
// its blur taps span, at most ceil(s*(UW-1)) + 5 by ceil(s*(UH-1)) + 5.
template <int TW, int TH, int SNUM>
struct PyrCap {
  static constexpr int UW = TW / 2 + 4, UH = TH / 2 + 4;
  static constexpr int RW = (SNUM * (UW - 1) + 3) / 4 + 5, RH = (SNUM * (UH - 1) + 3) / 4 + 5;
  HH_HD static bool fits(const PyrRegion& r) {
    return r.uw <= UW && r.uh <= UH && r.rw <= RW && r.rh <= RH;
  }
};

struct PyrLevelArgs {
  const PyrImage* imgs;
  const PyrTile* tiles;
  int ntiles;
  int level;
  int boundary;
  float w[25];
};

}  // namespace hh
