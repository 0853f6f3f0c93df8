// hipacc_hip_red.hpp -- global operators of the MI355X host runtime:
// Kernel::reduced_data() and Kernel::binned_data().  Included by
// hipacc_hip.hpp.
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

// Counting histogram of an accessor window into num_bins bins (bin index
// (p * num_bins) >> 8 for 8-bit, >> 16 for 16-bit pixels, (uint)(p * scale)
// for float pixels; out-of-range indices dropped).  Returns a host array of
// num_bins counts the caller owns (delete[]), as binned_data() does
// (replaces hipaccApplyBinningSegmented, runtime/hipacc_cu.tpp:410-464).
template <typename T>
unsigned int* hipaccApplyBinning(const HipaccAccessor<T>& acc, unsigned int num_bins,
                                 float scale = 1.0f,
                                 const HipaccExecutionParameterHip& ep = nullptr) {
  unsigned int* result = new unsigned int[num_bins]();
  void* bins = nullptr;
  int32_t stride = 0;
  if (hipaccCheck(hipacc_hip_alloc2d(&bins, &stride, (int32_t)num_bins, 1, 4, 4),
                  "hipaccApplyBinning"))
    return result;
  hipacc_view v = hipaccView(acc);
  if (ep) ep->pre_kernel();
  if (!hipaccCheck(hipacc_hip_histogram(&v, (int32_t)num_bins, scale, (uint32_t*)bins,
                                        hipaccStreamOf(ep)),
                   "hipaccApplyBinning"))
    hipaccCheck(hipacc_hip_download2d(result, bins, (int32_t)num_bins, (int32_t)num_bins, 1, 4,
                                      hipaccStreamOf(ep)),
                "hipaccApplyBinning");
  if (ep) ep->post_kernel();
  hipaccCheck(hipacc_hip_free(bins), "hipaccApplyBinning");
  return result;
}

#endif  // HIPACC_HIP_RED_HPP
