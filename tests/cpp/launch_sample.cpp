// A host program that launches a kernel of its own the way the reference's
// generated host code does: hipacc_launch_info for the iteration space,
// hipaccCalcGridFromBlock, hipaccPrepareKernelLaunch and hipaccLaunchKernel
// on an execution parameter's stream.  The kernel is a 3x3 mean in the
// generated kernels' form: blocks inside the border bounds read without
// boundary handling, the others clamp.  Interior reads are still clamped to
// the image, and a counter records any interior read that needed it (a
// wrong border bound), so the check cannot read outside the image.
// Several image sizes, offsets and block shapes; prints PASSED when every
// output equals the host's 3x3 clamped mean, no interior read left the
// image, the grid covers the iteration space and the timed launch reported
// a time.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <vector>

#include "hipacc_hip_device.hpp"

// Output pixel (x, y) of the iteration space [ox, ox + isw) x [oy, oy + ish)
// of a w x h image, the grid laid out as the generated kernels lay it out:
// columns from image column 0 (threads left of ox idle), rows from oy.
__global__ void mean3(const float* in, float* out, int w, int h, int stride, int ox, int oy,
                      int isw, int ish, int bh_l, int bh_r, int bh_t, int bh_b, int fall_back,
                      unsigned* violations) {
  const int x = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const int y = oy + (int)(blockIdx.y * blockDim.y + threadIdx.y);
  if (x < ox || x >= ox + isw || y >= oy + ish) return;
  const bool interior = !fall_back && (int)blockIdx.x >= bh_l && (int)blockIdx.x < bh_r &&
                        (int)blockIdx.y >= bh_t && (int)blockIdx.y < bh_b;
  float s = 0.0f;
  bool left_image = false;
  for (int dy = -1; dy <= 1; ++dy)
    for (int dx = -1; dx <= 1; ++dx) {
      const int xx = x + dx, yy = y + dy;
      const int cx = min(max(xx, 0), w - 1), cy = min(max(yy, 0), h - 1);
      left_image |= cx != xx || cy != yy;
      s += in[(long long)cy * stride + cx];
    }
  if (interior && left_image) atomicAdd(violations, 1u);
  out[(long long)y * stride + x] = s / 9.0f;
}

static int run(int w, int h, int ox, int oy, int bx, int by) {
  std::vector<float> host((size_t)w * h);
  for (int y = 0; y < h; ++y)
    for (int x = 0; x < w; ++x) host[(size_t)y * w + x] = (float)((x * 7 + y * 13) % 31);
  HipaccImageHip<float> in = hipaccCreateMemory<float>(host.data(), w, h);
  HipaccImageHip<float> out = hipaccCreateMemory<float>(nullptr, w, h);
  const int isw = w - 2 * ox, ish = h - 2 * oy;
  HipaccAccessor<float> is = hipaccMakeAccessor(out, isw, ish, ox, oy);
  hipacc_launch_info info(1, 1, is, 1, 1);   // a 3x3 window: half size 1
  const dim3 block(bx, by);
  const dim3 grid = hipaccCalcGridFromBlock(info, block);
  hipaccPrepareKernelLaunch(info, block);
  if ((int)(grid.x * block.x) < ox + isw || (int)(grid.y * block.y) < ish) {
    std::printf("grid %u x %u does not cover %d x %d\n", grid.x, grid.y, ox + isw, ish);
    return 1;
  }
  unsigned* violations = nullptr;
  int32_t vstride = 0;
  if (hipaccCheck(hipacc_hip_alloc2d((void**)&violations, &vstride, 1, 1, 4, 0), "alloc")) return 1;
  hipaccCheck(hipacc_hip_fill(violations, 0u, 1, 4, nullptr), "fill");
  hipaccSetTiming(false);
  HipaccExecutionParameterHip ep = nullptr;
  hipaccLaunchKernel(mean3, grid, block, ep, true, 0, in->get_device_memory(),
                     out->get_device_memory(), w, h, in->get_stride(), ox, oy, isw, ish,
                     info.bh_start_left, info.bh_start_right, info.bh_start_top,
                     info.bh_start_bottom, info.bh_fall_back, violations);
  const float ms = hipacc_last_kernel_timing();
  unsigned v = 0;
  hipaccCheck(hipacc_hip_download2d(&v, violations, 1, 1, 1, 4, nullptr), "download");
  (void)hipacc_hip_free(violations);
  const float* res = hipaccReadMemory<float>(out);
  int bad = 0;
  for (int y = oy; y < oy + ish; ++y)
    for (int x = ox; x < ox + isw; ++x) {
      float s = 0.0f;
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          const int cx = std::min(std::max(x + dx, 0), w - 1);
          const int cy = std::min(std::max(y + dy, 0), h - 1);
          s += host[(size_t)cy * w + cx];
        }
      if (res[(size_t)y * w + x] != s / 9.0f) ++bad;
    }
  std::printf("%dx%d window (%d,%d) %dx%d block %dx%d: grid %ux%u, interior blocks x [%d,%d)"
              " y [%d,%d) fall_back %d, %d wrong, %u interior reads outside, %.4f ms\n",
              w, h, ox, oy, isw, ish, bx, by, grid.x, grid.y, info.bh_start_left,
              info.bh_start_right, info.bh_start_top, info.bh_start_bottom, info.bh_fall_back,
              bad, v, ms);
  return bad || v || !(ms > 0.0f);
}

int main() {
  hipaccInitHIP(0);
  int fails = 0;
  fails += run(1000, 700, 0, 0, 32, 4);
  fails += run(1000, 700, 5, 3, 64, 2);
  fails += run(129, 65, 2, 1, 16, 8);
  fails += run(40, 9, 0, 0, 32, 4);       // too small for interior blocks: fall-back
  std::printf(fails ? "FAILED\n" : "PASSED\n");
  return fails ? 1 : 0;
}
