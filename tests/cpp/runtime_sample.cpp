// runtime_sample.cpp -- a host program written against the C++ runtime
// (hipacc_amd/runtime/hipacc_hip.hpp) the way generated host code uses the
// reference runtime: images, accessor windows, masks, local / separable /
// point operators, global reduce and binning, region copies and a pyramid
// traversal.  Writes every result to <outdir>/<name>.bin (dense rows) and
// the scalars to <outdir>/scalars.txt; the GPU test checks them against the
// oracle.
//
// Usage: runtime_sample <outdir>
#include <cstdio>
#include <fstream>
#include <string>
#include <vector>

#include "../../hipacc_amd/runtime/hipacc_hip.hpp"

static std::string g_out;

template <typename T>
static void save(const std::string& name, const T* data, size_t n) {
  std::ofstream f(g_out + "/" + name + ".bin", std::ios::binary);
  f.write(reinterpret_cast<const char*>(data), (std::streamsize)(n * sizeof(T)));
}

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s <outdir>\n", argv[0]);
    return 2;
  }
  g_out = argv[1];
  hipaccInitHIP(0);

  const int W = 203, H = 117;
  std::vector<unsigned char> src8((size_t)W * H);
  std::vector<float> srcf((size_t)W * H);
  for (int y = 0; y < H; ++y)
    for (int x = 0; x < W; ++x) {
      unsigned v = (unsigned)(x * 37 + y * 91 + (x * y) % 13) & 255u;
      src8[(size_t)y * W + x] = (unsigned char)v;
      srcf[(size_t)y * W + x] = (float)v / 255.0f;
    }

  // Gaussian 5x5 on uint8, rounded output, pitched images (256-byte rows).
  const float g5[25] = {
      0.00296902f, 0.01330621f, 0.02193823f, 0.01330621f, 0.00296902f,
      0.01330621f, 0.05963430f, 0.09832033f, 0.05963430f, 0.01330621f,
      0.02193823f, 0.09832033f, 0.16210282f, 0.09832033f, 0.02193823f,
      0.01330621f, 0.05963430f, 0.09832033f, 0.05963430f, 0.01330621f,
      0.00296902f, 0.01330621f, 0.02193823f, 0.01330621f, 0.00296902f};
  HipaccImageHip<unsigned char> in8 = hipaccCreateMemory<unsigned char>(src8.data(), W, H, 256);
  HipaccImageHip<unsigned char> gauss = hipaccCreateMemory<unsigned char>(nullptr, W, H, 256);
  HipaccMask<float> gmask(5, 5, g5);
  HipaccLocalOptions gopt;
  gopt.boundary = HipaccBoundary::CLAMP;
  gopt.round = true;
  int st = hipaccLaunchLocal(hipaccMakeAccessor(in8), hipaccMakeAccessor(gauss), gmask, gopt);
  save("gauss_u8", hipaccReadMemory(gauss), (size_t)W * H);

  // Kernel timing (hipacc_last_kernel_timing): one timed launch, then one
  // with timing off, which leaves the last time as it was.
  hipaccSetTiming(true);
  st |= hipaccLaunchLocal(hipaccMakeAccessor(in8), hipaccMakeAccessor(gauss), gmask, gopt);
  const float t_on = hipacc_last_kernel_timing();
  hipaccSetTiming(false);
  st |= hipaccLaunchLocal(hipaccMakeAccessor(in8), hipaccMakeAccessor(gauss), gmask, gopt);
  const float t_off = hipacc_last_kernel_timing();

  // Sobel-x with a Domain from the mask (holes at the zero column), int
  // accumulation, int16 output, MIRROR.
  const int sx[9] = {-1, 0, 1, -2, 0, 2, -1, 0, 1};
  HipaccMask<int> smask(3, 3);
  hipaccWriteDomainFromMask(smask, sx);
  HipaccImageHip<short> sob = hipaccCreateMemory<short>(nullptr, W, H);
  HipaccLocalOptions sopt;
  sopt.boundary = HipaccBoundary::MIRROR;
  st |= hipaccLaunchLocal(hipaccMakeAccessor(in8), hipaccMakeAccessor(sob), smask, sopt);
  save("sobel_i16", hipaccReadMemory(sob), (size_t)W * H);

  // Separable 7-tap on float, REPEAT, on an explicit stream parameter.
  const std::vector<float> k7 = {0.0366328f, 0.111281f, 0.216745f, 0.270682f,
                                 0.216745f, 0.111281f, 0.0366328f};
  HipaccImageHip<float> inf = hipaccCreateMemory<float>(srcf.data(), W, H, 256);
  HipaccImageHip<float> sep = hipaccCreateMemory<float>(nullptr, W, H, 256);
  HipaccExecutionParameterHip ep = std::make_shared<HipaccStreamParameter>(nullptr);
  HipaccLocalOptions popt;
  popt.boundary = HipaccBoundary::REPEAT;
  st |= hipaccLaunchSeparable(hipaccMakeAccessor(inf), hipaccMakeAccessor(sep), k7, k7, popt, ep);
  save("sep_f32", hipaccReadMemory(sep), (size_t)W * H);

  // Point operator: |in - sep|.
  HipaccImageHip<float> diff = hipaccCreateMemory<float>(nullptr, W, H);
  HipaccAccessor<float> acc_sep = hipaccMakeAccessor(sep);
  st |= hipaccLaunchPoint(HIPACC_POINT_ABSDIFF, hipaccMakeAccessor(inf), hipaccMakeAccessor(diff),
                          &acc_sep);
  save("absdiff_f32", hipaccReadMemory(diff), (size_t)W * H);

  // Accessor windows: a 3x3 box MAX over the window (5,7,100,50) of the
  // float image into the window (3,2,100,50) of a 120x60 image.
  HipaccImageHip<float> win = hipaccCreateMemory<float>(nullptr, 120, 60);
  HipaccMask<float> box3(3, 3);
  HipaccLocalOptions mopt;
  mopt.reduce = HipaccReduce::MAX;
  mopt.use_weights = false;
  mopt.boundary = HipaccBoundary::CONSTANT;
  st |= hipaccLaunchLocal(hipaccMakeAccessor(inf, 100, 50, 5, 7),
                          hipaccMakeAccessor(win, 100, 50, 3, 2), box3, mopt);
  save("window_max_f32", hipaccReadMemory(win), (size_t)120 * 60);

  // Region copy.
  HipaccImageHip<unsigned char> cpy = hipaccCreateMemory<unsigned char>(nullptr, 64, 32);
  hipaccCopyMemoryRegion(hipaccMakeAccessor(in8, 40, 20, 11, 13),
                         hipaccMakeAccessor(cpy, 40, 20, 4, 6));
  save("copy_u8", hipaccReadMemory(cpy), (size_t)64 * 32);

  // Global operators.
  float fsum = hipaccApplyReduction<float>(inf, HipaccReduce::SUM);
  float fmin = hipaccApplyReduction<float>(inf, HipaccReduce::MIN);
  float fmax = hipaccApplyReduction<float>(hipaccMakeAccessor(inf, 50, 40, 20, 30),
                                           HipaccReduce::MAX);
  unsigned int usum = hipaccApplyReduction<unsigned int>(in8, HipaccReduce::SUM);
  unsigned int* hist = hipaccApplyBinning(hipaccMakeAccessor(in8), 256);
  save("hist_u32", hist, 256);
  delete[] hist;
  // the same through the names generated host code calls
  float fmax_shared = hipaccApplyReductionShared<float>(
      HipaccReduce::MAX, hipaccMakeAccessor(inf, 50, 40, 20, 30), 256, 16, ep);
  unsigned int* hist2 = hipaccApplyBinningSegmented<unsigned int, unsigned char>(
      HipaccCountBins{}, hipaccMakeAccessor(in8), 4, 32, 256, ep);
  save("hist2_u32", hist2, 256);
  delete[] hist2;
  // repeated binning reuses the thread's device bins
  for (int i = 0; i < 3; ++i) delete[] hipaccApplyBinning(hipaccMakeAccessor(in8), 64);

  // Pyramid: depth 4; at each level below the top, the level image is the
  // top-left window of the level above (a copy), and a traversal counts the
  // levels it visits.
  HipaccPyramidHip<float> pyr = hipaccCreatePyramid<float>(inf, 4);
  HipaccPyramidTraversor trav;
  int visited = 0;
  trav.hipaccTraverse(pyr, [&] {
    ++visited;
    if (!pyr.is_top_level()) {
      HipaccImageHip<float> lvl = pyr(0), up = pyr(-1);
      hipaccCopyMemoryRegion(hipaccMakeAccessor(up, lvl->get_width(), lvl->get_height(), 0, 0),
                             hipaccMakeAccessor(lvl));
    }
    trav.hipaccTraverse();
  });
  HipaccImageHip<float> last = pyr(3);
  save("pyr3_f32", hipaccReadMemory(last), (size_t)last->get_width() * last->get_height());

  std::FILE* f = std::fopen((g_out + "/scalars.txt").c_str(), "w");
  if (!f) return 1;
  std::fprintf(f, "status %d\nfsum %.9g\nfmin %.9g\nfmax %.9g\nusum %u\nvisited %d\n"
               "pyr3 %d %d\nstride_gauss %d\nstride_sobel %d\ntiming %.9g %.9g\n"
               "fmax_shared %.9g\n",
               st, fsum, fmin, fmax, usum, visited, last->get_width(), last->get_height(),
               gauss->get_stride(), sob->get_stride(), t_on, t_off, fmax_shared);
  std::fclose(f);
  return st == 0 ? 0 : 1;
}
