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

struct ResampleArgs {
  const float* in;
  const float* in2;
  float* out;
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
struct PyrLevelArgs {
  const PyrImage* imgs;
  const PyrTile* tiles;
  int ntiles;
  int level;
  int boundary;
  float w[25];
};

}  // namespace hh
