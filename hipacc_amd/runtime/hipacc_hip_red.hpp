// hipacc_hip_red.hpp -- global operators of the MI355X host runtime:
// Kernel::reduced_data() and Kernel::binned_data().  Included by
// hipacc_hip.hpp.
//
// Two spellings of each: hipaccApplyReduction / hipaccApplyBinning, and the
// names the reference's generated host code calls,
// hipaccApplyReductionShared / hipaccApplyBinningSegmented
// (lib/Rewrite/CreateHostStrings.cpp:1033,1151; runtime/hipacc_cu.tpp:
// 311-464), with the same parameter order.  Where generated code passes its
// generated reduction or binning kernel, a HIP emitter passes the operator
// (HipaccReduce mode, HipaccCountBins); the block-shape parameters are
// accepted and ignored, since the library chooses its own launch shape.
#ifndef HIPACC_HIP_RED_HPP
#define HIPACC_HIP_RED_HPP

// Reduces an accessor window with `reduce` in accumulator type ACC (float,
// int or unsigned int) and returns the value; blocking, as reduced_data()
// (replaces hipaccApplyReductionShared, runtime/hipacc_cu.tpp:311-408).
// Float SUM is a fixed-shape tree: deterministic, equal to the sequential
// sum only within rounding.
template <typename ACC, typename T>
ACC hipaccApplyReduction(const HipaccAccessor<T>& acc, HipaccReduce reduce,
                         const HipaccExecutionParameterHip& ep = nullptr) {
  hipacc_view v = hipaccView(acc);
  ACC result{};
  if (ep) ep->pre_kernel();
  hipaccCheck(hipacc_hip_reduce_sync(&v, (int32_t)reduce, hipacc_pixel_code<ACC>::value, &result,
                                     hipaccStreamOf(ep)),
              "hipaccApplyReduction");
  if (ep) ep->post_kernel();
  return result;
}
template <typename ACC, typename T>
ACC hipaccApplyReduction(const HipaccImageHip<T>& img, HipaccReduce reduce,
                         const HipaccExecutionParameterHip& ep = nullptr) {
  return hipaccApplyReduction<ACC>(HipaccAccessor<T>(img), reduce, ep);
}

// The generated-code spelling: hipaccApplyReductionShared<T>(op, acc,
// max_threads, pixels_per_thread, ep, tex, print_timing); the result type is
// the accessor's pixel type, as there.  print_timing times the reduction and
// logs it, as the reference's does.
template <typename T>
T hipaccApplyReductionShared(HipaccReduce reduce, const HipaccAccessor<T>& acc,
                             unsigned int /*max_threads*/, unsigned int /*pixels_per_thread*/,
                             const HipaccExecutionParameterHip& ep,
                             const void* /*tex*/ = nullptr, bool print_timing = false) {
  if (print_timing) hipaccSetTiming(true);
  const T r = hipaccApplyReduction<T>(acc, reduce, ep);
  if (print_timing) {
    hipaccSetTiming(false);
    hipaccRuntimeLogTrivial(hipaccRuntimeLogLevel::INFO,
                            "<HIPACC:> Reduction timing: " +
                                std::to_string(hipacc_last_kernel_timing()) + "(ms)");
  }
  return r;
}
template <typename T>
T hipaccApplyReductionShared(HipaccReduce reduce, const HipaccImageHip<T>& img,
                             unsigned int max_threads, unsigned int pixels_per_thread,
                             const HipaccExecutionParameterHip& ep,
                             const void* tex = nullptr, bool print_timing = false) {
  return hipaccApplyReductionShared<T>(reduce, HipaccAccessor<T>(img), max_threads,
                                       pixels_per_thread, ep, tex, print_timing);
}

// Device bins of the calling host thread, grown on demand and freed when the
// thread ends: binning allocates nothing per call once warm.
struct HipaccBinBuffer {
  void* bins = nullptr;
  unsigned int capacity = 0;
  ~HipaccBinBuffer() {
    if (bins) (void)hipacc_hip_free(bins);
  }
  uint32_t* get(unsigned int n) {
    if (n > capacity) {
      if (bins) hipaccCheck(hipacc_hip_free(bins), "hipaccApplyBinning");
      bins = nullptr;
      capacity = 0;
      int32_t stride = 0;
      if (hipaccCheck(hipacc_hip_alloc2d(&bins, &stride, (int32_t)n, 1, 4, 4),
                      "hipaccApplyBinning"))
        return nullptr;
      capacity = n;
    }
    return static_cast<uint32_t*>(bins);
  }
};
inline HipaccBinBuffer& hipaccBinBuffer() {
  static thread_local HipaccBinBuffer buf;
  return buf;
}

// Counting histogram of an accessor window into num_bins bins (bin index
// (p * num_bins) >> 8 for 8-bit, >> 16 for 16-bit pixels, (uint)(p * scale)
// for float pixels; out-of-range indices dropped).  Returns a host array of
// num_bins counts the caller owns (delete[]), as binned_data() does
// (replaces hipaccApplyBinningSegmented, runtime/hipacc_cu.tpp:410-464).
// Blocking.
template <typename T>
unsigned int* hipaccApplyBinning(const HipaccAccessor<T>& acc, unsigned int num_bins,
                                 float scale = 1.0f,
                                 const HipaccExecutionParameterHip& ep = nullptr) {
  unsigned int* result = new unsigned int[num_bins]();
  uint32_t* bins = hipaccBinBuffer().get(num_bins);
  if (!bins) return result;
  hipacc_view v = hipaccView(acc);
  if (ep) ep->pre_kernel();
  if (!hipaccCheck(hipacc_hip_histogram(&v, (int32_t)num_bins, scale, bins, hipaccStreamOf(ep)),
                   "hipaccApplyBinning"))
    hipaccCheck(hipacc_hip_download2d(result, bins, (int32_t)num_bins, (int32_t)num_bins, 1, 4,
                                      hipaccStreamOf(ep)),
                "hipaccApplyBinning");
  if (ep) ep->post_kernel();
  return result;
}

// The binning operator a HIP emitter passes where generated code passes its
// binning kernel: a counting histogram (bin(idx) = 1, reduce a + b) with the
// float pixels' bin scale.
struct HipaccCountBins {
  float scale = 1.0f;
};

// The generated-code spelling: hipaccApplyBinningSegmented<Bin, Pix>(op, acc,
// num_warps, num_units, num_bins, ep, tex, print_timing), returning a new
// Bin[num_bins] the caller owns.
template <typename Bin, typename Pix>
Bin* hipaccApplyBinningSegmented(const HipaccCountBins& op, const HipaccAccessor<Pix>& acc,
                                 unsigned int /*num_warps*/, unsigned int /*num_units*/,
                                 unsigned int num_bins, const HipaccExecutionParameterHip& ep,
                                 const void* /*tex*/ = nullptr, bool print_timing = false) {
  if (print_timing) hipaccSetTiming(true);
  unsigned int* r = hipaccApplyBinning(acc, num_bins, op.scale, ep);
  if (print_timing) {
    hipaccSetTiming(false);
    hipaccRuntimeLogTrivial(hipaccRuntimeLogLevel::INFO,
                            "<HIPACC:> Binning timing: " +
                                std::to_string(hipacc_last_kernel_timing()) + "(ms)");
  }
  Bin* out = new Bin[num_bins];
  for (unsigned int i = 0; i < num_bins; ++i) out[i] = (Bin)r[i];
  delete[] r;
  return out;
}

#endif  // HIPACC_HIP_RED_HPP
