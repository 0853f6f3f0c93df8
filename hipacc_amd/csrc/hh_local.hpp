// hh_local.hpp -- argument blocks shared by the local-operator kernels and the
// host dispatch code.
#pragma once

#include "hh_common.hpp"

namespace hh {

constexpr int kMaxTaps = HIPACC_MAX_MASK_DIM * HIPACC_MAX_MASK_DIM;

// Geometry of one accessor / iteration-space pair, resolved on the host.
struct Geom {
  const void* in;        // image 0, pixel (0,0)
  void* out;
  long long in_plane, out_plane;  // elements between batch images
  int in_stride, out_stride;      // elements
  int in_dtype, out_dtype;
  int in_imgw, in_imgh;           // allocation extent of the input
  int lo_x, lo_y, hi_x, hi_y;     // accessor window [lo, hi)
  int in_ox, in_oy;               // accessor offset (== lo)
  int out_ox, out_oy;             // iteration-space offset
  int is_w, is_h;                 // iteration-space extent
  int batch;                      // images (grid z)
  int boundary;
  float const_f;                  // CONSTANT value (float view)
  int const_i;                    // CONSTANT value (integer view)
  int cvt;
  float post_div;
};

union Weights {
  float f[kMaxTaps];
  int32_t i[kMaxTaps];
};

struct LocalArgs {
  Geom g;
  int mask_w, mask_h;
  int reduce;
  int use_weights;
  int fma;          // fp_contract
  int in_vec_ok;    // input window admits aligned 16-byte row loads
  int out_vec_ok;   // output window admits aligned vector stores
  const float* wdev;
  uint8_t domain[kMaxTaps];
  Weights w;
};

// Kernel variants of a local operator (the autotuner's choice):
//  tile    LDS-staged tile + halo, PPT rows per wave (hh_local_fast.hpp)
//  stream  row-streaming strips, `param` rows per segment (hh_stream.hpp)
//  generic one pixel per thread, any shape / mode (local.hip)
//  wide    stream with 8 pixels per lane instead of 4
//  ring    block row ring in LDS, `param` rows per block segment (hh_ring.hpp)
enum Variant {
  kVarAuto = 0, kVarTile = 1, kVarStream = 2, kVarGeneric = 3, kVarStreamWide = 4, kVarRing = 5
};

// Operator kinds the tuner keys on.
enum TuneKind { kTuneLocal = 0, kTuneSeparable = 1, kTuneBox = 2 };
// sz: size class of the iteration space, floor(log2(width * height *
// batch)); a variant tuned at one size is not replayed at another.
struct TuneKey {
  int kind, kx, ky, in_dt, out_dt, acc_f, boundary, fma, wdev, sz;
};
inline int tune_size_class(long long w, long long h, long long batch) {
  unsigned long long n = (unsigned long long)(w * h * (batch > 1 ? batch : 1));
  int c = 0;
  while (n > 1) {
    n >>= 1;
    ++c;
  }
  return c;
}
struct TuneVal {
  int variant, param;
  float ms;
};
bool tune_lookup(const TuneKey& k, TuneVal& v);
void tune_store(const TuneKey& k, const TuneVal& v);
int tune_save(const char* path);
int tune_load(const char* path);

// Launch one local operator; returns a HIPACC_* status.  variant/param
// force a kernel variant (kVarAuto: the tuned or default choice).
int launch_local(const LocalArgs& a, int acc_is_float, hipStream_t stream, int variant, int param);
// Local operators on 4-wide vector pixels (local_vec4.hip).
int launch_local_vec4(const LocalArgs& a, int acc_is_float, hipStream_t s);
// Fast weighted-SUM instantiations, one table per pixel-type family; each
// returns HIPACC_ERR_UNSUPPORTED when it has no matching instantiation.
int dispatch_sum_f32(const LocalArgs& a, int ppt, hipStream_t s);
int dispatch_sum_u8(const LocalArgs& a, int ppt, hipStream_t s);
int dispatch_sum_int(const LocalArgs& a, int ppt, hipStream_t s);

struct SepArgs {
  Geom g;
  int kx, ky;
  int in_vec_ok, out_vec_ok;
  float rowf[2 * HIPACC_MAX_SEP_RADIUS + 1];
  float colf[2 * HIPACC_MAX_SEP_RADIUS + 1];
  int32_t rowi[2 * HIPACC_MAX_SEP_RADIUS + 1];
  int32_t coli[2 * HIPACC_MAX_SEP_RADIUS + 1];
  int box;          // integer weights are all 1: running sums are exact
};
int launch_separable(const SepArgs& a, int acc_is_float, int fp_contract,
                     hipStream_t stream, int variant, int param);
int dispatch_stream_local(const LocalArgs& a, int acc_is_float, int seg, int wide, hipStream_t s);
int dispatch_ring_local(const LocalArgs& a, int seg_rows, hipStream_t s);
int dispatch_ring8_local(const LocalArgs& a, int acc_is_float, int seg_rows, hipStream_t s);
bool ring8_pays(const LocalArgs& a);
int dispatch_ring_sep(const SepArgs& a, int acc_is_float, int fma, int seg_rows, hipStream_t s);
int dispatch_stream_sep(const SepArgs& a, int acc_is_float, int fma, int seg, int wide,
                        hipStream_t s);

// ---- point operators ------------------------------------------------------
struct PointArgs {
  const void* in0;
  const void* in1;
  void* out;
  long long plane0, plane1, plane_out;
  int stride0, stride1, stride_out;
  int ox0, oy0, ox1, oy1, oxo, oyo;
  int w, h, batch;
  int op, cvt;
  float alpha, beta;
  int vec_ok;
};
int launch_point(const PointArgs& a, int dt0, int dt1, int dto, hipStream_t s);

// ---- global reduction -----------------------------------------------------
constexpr int kRedBlocks = 1024;
struct ReduceArgs {
  const void* in;
  int stride, ox, oy, w, h;
  int mode;
  int vec_ok;
  void* result;
  unsigned long long* partials;   // kRedBlocks entries
};
int launch_reduce(const ReduceArgs& a, int dt, int acc_dt, hipStream_t s);

// ---- histogram --------------------------------------------------------------
struct HistArgs {
  const void* in;
  int stride, ox, oy, w, h;
  int nbins;
  float scale;
  int vec_ok;
  uint32_t* bins;
  uint32_t* partials;   // [blocks][nbins] per-block histograms (workspace)
};
// Entries of the histogram's partial-bin workspace: blocks * nbins <= this.
constexpr int kHistPartialEntries = 1 << 18;
int launch_histogram(const HistArgs& a, int dt, hipStream_t s);

// ---- utilities --------------------------------------------------------------
int launch_fill(void* p, uint32_t value, size_t n, int eb, hipStream_t s);
int launch_random(void* p, int dt, int stride, int ox, int oy, int w, int h, int batch,
                  long long plane, uint64_t seed, hipStream_t s);

// ---- bilateral --------------------------------------------------------------
struct BilatArgs {
  Geom g;
  int ksize;
  int in_vec_ok, out_vec_ok;
  float k2;                 // -c_r * log2(e)
  float lsp[kMaxTaps];      // log2(spatial)
  float* norm;              // F32 out only: per-pixel SUM of weights, or NULL
  long long norm_plane;
  int norm_stride;
};
int launch_bilateral(const BilatArgs& a, hipStream_t s);

// Records the detailed message of the calling thread's last failure.
int set_error(int status, const char* fmt, ...);
int hip_status(hipError_t e, const char* what);

}  // namespace hh
