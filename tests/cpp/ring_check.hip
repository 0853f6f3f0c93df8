// ring_check.hip -- host-side check of the halo-ring and border-band
// enumerations of the adjoint kernels (hipacc_amd/csrc/hh_grad.hpp): every
// extended position outside the image appears once in the ring, ring_index
// inverts ring_pos, and the band is exactly the pixels within (rx, ry) of an
// edge, each once.  Prints "ok" or the first failure.
#include <cstdio>
#include <set>
#include <utility>

#include "../../hipacc_amd/csrc/hh_grad.hpp"

int main() {
  const int sizes[][2] = {{1, 1}, {2, 3}, {3, 2}, {5, 7}, {8, 8}, {9, 4}, {37, 29}, {64, 16}, {130, 3}};
  const int radii[][2] = {{1, 1}, {2, 2}, {4, 4}, {1, 3}, {3, 1}, {7, 7}};
  for (auto& sz : sizes)
    for (auto& rr : radii) {
      const int w = sz[0], h = sz[1], rx = rr[0], ry = rr[1];
      std::set<std::pair<int, int>> seen;
      const long long n = hh::ring_size(w, h, rx, ry);
      for (long long j = 0; j < n; ++j) {
        int ix, iy;
        hh::ring_pos(w, h, rx, ry, j, ix, iy);
        const bool inside = ix >= 0 && ix < w && iy >= 0 && iy < h;
        const bool in_ext = ix >= -rx && ix < w + rx && iy >= -ry && iy < h + ry;
        if (inside || !in_ext || !seen.insert({ix, iy}).second ||
            hh::ring_index(w, h, rx, ry, ix, iy) != j) {
          std::printf("ring fail w=%d h=%d rx=%d ry=%d j=%lld (%d,%d)\n", w, h, rx, ry, j, ix, iy);
          return 1;
        }
      }
      const long long ext = (long long)(w + 2 * rx) * (h + 2 * ry) - (long long)w * h;
      if ((long long)seen.size() != ext) {
        std::printf("ring count w=%d h=%d: %zu of %lld\n", w, h, seen.size(), ext);
        return 1;
      }
      std::set<std::pair<int, int>> band;
      const long long nb = hh::band_size(w, h, rx, ry);
      for (long long j = 0; j < nb; ++j) {
        int x, y;
        hh::band_pos(w, h, rx, ry, j, x, y);
        const bool near = x < rx || x >= w - rx || y < ry || y >= h - ry;
        if (x < 0 || x >= w || y < 0 || y >= h || !near || !band.insert({x, y}).second) {
          std::printf("band fail w=%d h=%d rx=%d ry=%d j=%lld (%d,%d)\n", w, h, rx, ry, j, x, y);
          return 1;
        }
      }
      long long want = 0;
      for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) want += (x < rx || x >= w - rx || y < ry || y >= h - ry);
      if (nb != want) {
        std::printf("band count w=%d h=%d rx=%d ry=%d: %lld of %lld\n", w, h, rx, ry, nb, want);
        return 1;
      }
    }
  std::printf("ok\n");
  return 0;
}
