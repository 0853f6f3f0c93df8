/*
 * hipacc_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * A CPU restatement of what Hipacc computes for its data-parallel operators,
 * used by the tests, by __graft_entry__.smoke() and by bench.py's CPU
 * baseline as the checker.  Nothing in the product (libhipacc_gfx950.so,
 * the hipacc_amd package) links, loads or calls this file.
 *
 * Each oracle_* function takes the same descriptor structs as the C ABI in
 * include/hipacc_hip.h, but every `data` pointer is HOST memory.  The loop
 * structure follows the reference's generated CPU backend: rows in parallel
 * with `#pragma omp parallel for` (lib/Backend/CPU_x86.cpp:3670-3689), every
 * pixel with boundary handling (lib/AST/ASTTranslate.cpp:207-224), taps in
 * row-major order (lib/AST/Convolution.cpp:397-435; dsl/kernel.hpp:241-296),
 * reductions and binning partitioned as runtime/hipacc_cpu_red.hpp:19-126.
 *
 * Built with -ffp-contract=off so that `a + w*x` rounds twice, as the DSL
 * host emulation compiled for x86-64 without FMA does; fp_contract=1 uses
 * C99 fmaf(), which is exactly the fused operation the GPU path issues.
 *
 * Every function is pinned by checks outside this file: the golden vectors
 * and known answers of the reference's own DSL programs (tests/golden/), and
 * closed forms, library routines and brute force on small inputs
 * (tests/test_cpu_oracle_math.py, tests/test_cpu_oracle_grad.py).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/hipacc_hip.h"
#include "splitmix.h"

#define ORC_EXPORT __attribute__((visibility("default")))

/* ---- pixel access ------------------------------------------------------ */
/* Vector pixels (uchar4, short4, int4, float4) are four interleaved
 * channels of their scalar type; channel c of pixel i is element 4*i + c. */
static int dt_base(int dt) {
  switch (dt) {
    case HIPACC_U8X4: return HIPACC_U8;
    case HIPACC_I16X4: return HIPACC_I16;
    case HIPACC_I32X4: return HIPACC_I32;
    case HIPACC_F32X4: return HIPACC_F32;
    default: return dt;
  }
}
static int dt_lanes(int dt) { return dt >= HIPACC_U8X4 ? 4 : 1; }
static int dt_bytes(int dt) {
  switch (dt_base(dt)) {
    case HIPACC_U8: case HIPACC_I8: return 1 * dt_lanes(dt);
    case HIPACC_U16: case HIPACC_I16: return 2 * dt_lanes(dt);
    default: return 4 * dt_lanes(dt);
  }
}
static double px_get(const void* p, int dt, long long i) {
  switch (dt) {
    case HIPACC_U8: return ((const uint8_t*)p)[i];
    case HIPACC_I8: return ((const int8_t*)p)[i];
    case HIPACC_U16: return ((const uint16_t*)p)[i];
    case HIPACC_I16: return ((const int16_t*)p)[i];
    case HIPACC_U32: return ((const uint32_t*)p)[i];
    case HIPACC_I32: return ((const int32_t*)p)[i];
    default: return ((const float*)p)[i];
  }
}
static float px_f(const void* p, int dt, long long i) {
  if (dt == HIPACC_F32) return ((const float*)p)[i];
  return (float)px_get(p, dt, i);
}
static int32_t px_i(const void* p, int dt, long long i) {
  switch (dt) {
    case HIPACC_U8: return ((const uint8_t*)p)[i];
    case HIPACC_I8: return ((const int8_t*)p)[i];
    case HIPACC_U16: return ((const uint16_t*)p)[i];
    case HIPACC_I16: return ((const int16_t*)p)[i];
    case HIPACC_U32: return (int32_t)((const uint32_t*)p)[i];
    case HIPACC_I32: return ((const int32_t*)p)[i];
    default: return (int32_t)((const float*)p)[i];
  }
}

/* ---- output conversion: the C cast of dsl/types.hpp:115-117 for every
 * value the output type can hold; outside that range (undefined in the
 * reference) truncation saturates to the uint32 range for uint outputs and
 * to the int32 range otherwise (NaN -> 0), then wraps to the output width,
 * as stated for hipacc_cvt in include/hipacc_hip.h. ----------------------- */
static int32_t f2i_sat(float v) {
  if (!(v == v)) return 0;
  if (v >= 2147483648.0f) return 2147483647;
  if (v <= -2147483648.0f) return (int32_t)0x80000000u;
  return (int32_t)v;
}
static uint32_t f2u_sat(float v) {
  if (!(v == v) || v <= 0.0f) return 0u;
  if (v >= 4294967296.0f) return 0xffffffffu;
  return (uint32_t)v;
}
static void px_put_f(void* p, int dt, long long i, float acc, int cvt, float post_div) {
  if (post_div != 0.0f) acc = acc / post_div;
  if (cvt == HIPACC_CVT_ROUND) acc = acc + 0.5f;
  if (dt == HIPACC_F32) { ((float*)p)[i] = acc; return; }
  const uint32_t u = (uint32_t)f2i_sat(acc);
  switch (dt) {
    case HIPACC_U8: ((uint8_t*)p)[i] = (uint8_t)u; break;
    case HIPACC_I8: ((int8_t*)p)[i] = (int8_t)(uint8_t)u; break;
    case HIPACC_U16: ((uint16_t*)p)[i] = (uint16_t)u; break;
    case HIPACC_I16: ((int16_t*)p)[i] = (int16_t)(uint16_t)u; break;
    case HIPACC_U32: ((uint32_t*)p)[i] = f2u_sat(acc); break;
    case HIPACC_I32: ((int32_t*)p)[i] = (int32_t)u; break;
  }
}
static void px_put_i(void* p, int dt, long long i, int32_t acc, int cvt, float post_div) {
  if (post_div != 0.0f) { px_put_f(p, dt, i, (float)acc / post_div, cvt, 0.0f); return; }
  switch (dt) {
    case HIPACC_U8: ((uint8_t*)p)[i] = (uint8_t)acc; break;
    case HIPACC_I8: ((int8_t*)p)[i] = (int8_t)acc; break;
    case HIPACC_U16: ((uint16_t*)p)[i] = (uint16_t)acc; break;
    case HIPACC_I16: ((int16_t*)p)[i] = (int16_t)acc; break;
    case HIPACC_U32: ((uint32_t*)p)[i] = (uint32_t)acc; break;
    case HIPACC_I32: ((int32_t*)p)[i] = acc; break;
    default: ((float*)p)[i] = (float)acc; break;
  }
}

/* ---- boundary handling: Accessor::pixel_bh (dsl/image.hpp:574-612) with
 * the generated REPEAT (lib/AST/BorderHandling.cpp:59-74, periodic) and
 * single-reflection MIRROR (:78-103); window [lo, hi).  Indices a single
 * reflection leaves outside the window, and UNDEFINED reads outside the
 * image, are undefined in the reference and clamped here. ---------------- */
static int bh_clamp(int i, int lo, int hi) { return i < lo ? lo : (i >= hi ? hi - 1 : i); }
static int bh_index(int mode, int i, int lo, int hi, int ext) {
  switch (mode) {
    case HIPACC_BOUNDARY_CLAMP: return bh_clamp(i, lo, hi);
    case HIPACC_BOUNDARY_REPEAT: {
      while (i < lo) i += hi - lo;
      while (i >= hi) i -= hi - lo;
      return i;
    }
    case HIPACC_BOUNDARY_MIRROR:
      if (i < lo) i = lo + (lo - i - 1);
      if (i >= hi) i = hi - (i + 1 - hi);
      return bh_clamp(i, lo, hi);
    default: return bh_clamp(i, 0, ext);
  }
}
/* Returns 1 and sets *idx when the tap reads the image, 0 for the CONSTANT
 * value. */
static int bh_tap(const hipacc_view* v, int mode, int x, int y, long long* idx) {
  const int lox = v->offset_x, hix = v->offset_x + v->width;
  const int loy = v->offset_y, hiy = v->offset_y + v->height;
  if (mode == HIPACC_BOUNDARY_CONSTANT) {
    if (x < lox || x >= hix || y < loy || y >= hiy) return 0;
  } else {
    x = bh_index(mode, x, lox, hix, v->img_width);
    y = bh_index(mode, y, loy, hiy, v->img_height);
  }
  *idx = (long long)y * v->stride + x;
  return 1;
}

static const char* plane(const hipacc_view* v, int b) {
  return (const char*)v->data + (long long)b * v->plane_stride * dt_bytes(v->dtype);
}
static int nbatch(const hipacc_view* v) { return v->batch > 1 ? v->batch : 1; }

/* constant value in the input pixel type, seen as float / int */
static float const_f(int dt, double c) {
  switch (dt) {
    case HIPACC_U8: return (float)(uint8_t)(long long)c;
    case HIPACC_I8: return (float)(int8_t)(long long)c;
    case HIPACC_U16: return (float)(uint16_t)(long long)c;
    case HIPACC_I16: return (float)(int16_t)(long long)c;
    case HIPACC_U32: return (float)(uint32_t)(long long)c;
    case HIPACC_I32: return (float)(int32_t)(long long)c;
    default: return (float)c;
  }
}
static int32_t const_i(int dt, double c) {
  switch (dt) {
    case HIPACC_U8: return (uint8_t)(long long)c;
    case HIPACC_I8: return (int8_t)(long long)c;
    case HIPACC_U16: return (uint16_t)(long long)c;
    case HIPACC_I16: return (int16_t)(long long)c;
    case HIPACC_U32: return (int32_t)(uint32_t)(long long)c;
    case HIPACC_I32: return (int32_t)(long long)c;
    default: return (int32_t)(float)c;
  }
}

/* ---- local operator ---------------------------------------------------- *
 * Kernel::convolve / Kernel::reduce(Domain) (dsl/kernel.hpp:241-296): the
 * first active tap seeds the result, later taps are combined in row-major
 * order.  MIN/MAX follow hipacc::math::min/max (dsl/math_functions.hpp:
 * 349-369) called as min(tap, result).                                     */
ORC_EXPORT int oracle_local(const hipacc_local_desc* d) {
  const hipacc_view* in = &d->in;
  const hipacc_view* out = &d->out;
  const int mw = d->mask_w, mh = d->mask_h, rx = mw / 2, ry = mh / 2;
  const int accf = d->acc_dtype == HIPACC_F32;
  /* vector pixels: the scalar operator on each channel */
  const int nch = dt_lanes(in->dtype);
  if (dt_lanes(out->dtype) != nch) return 1;
  const int ib = dt_base(in->dtype), ob = dt_base(out->dtype);
  const float cf = const_f(ib, d->const_val);
  const int32_t ci = const_i(ib, d->const_val);
  for (int b = 0; b < nbatch(out); ++b)
  for (int ch = 0; ch < nch; ++ch) {
    const char* ip = plane(in, b);
    char* op = (char*)plane(out, b);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < out->height; ++y) {
      for (int x = 0; x < out->width; ++x) {
        const int cx = x + in->offset_x, cy = y + in->offset_y;
        float af = 0.0f;
        int32_t ai = 0;
        int first = 1;
        for (int dy = 0; dy < mh; ++dy) {
          for (int dx = 0; dx < mw; ++dx) {
            const int t = dy * mw + dx;
            if (d->domain && !d->domain[t]) continue;
            long long idx = 0;
            const int hit = bh_tap(in, d->boundary, cx + dx - rx, cy + dy - ry, &idx);
            if (accf) {
              const float v = hit ? px_f(ip, ib, idx * nch + ch) : cf;
              float w = 1.0f;
              if (d->use_weights) {
                if (d->weights_f32) w = d->weights_f32[t];
                else if (d->weights_i32) w = (float)d->weights_i32[t];
              }
              if (d->reduce == HIPACC_REDUCE_SUM && d->use_weights) {
                if (first) af = w * v;
                else if (d->fp_contract) af = fmaf(w, v, af);
                else { const float p = w * v; af = af + p; }
              } else {
                const float tap = d->use_weights ? w * v : v;
                if (first) af = tap;
                else switch (d->reduce) {
                  case HIPACC_REDUCE_MIN: af = tap < af ? tap : af; break;
                  case HIPACC_REDUCE_MAX: af = tap > af ? tap : af; break;
                  case HIPACC_REDUCE_PROD: af = af * tap; break;
                  default: af = af + tap; break;
                }
              }
            } else {
              const int32_t v = hit ? px_i(ip, ib, idx * nch + ch) : ci;
              int32_t w = 1;
              if (d->use_weights) {
                if (d->weights_i32) w = d->weights_i32[t];
                else if (d->weights_f32) w = (int32_t)d->weights_f32[t];
              }
              const int32_t tap = d->use_weights ? (int32_t)((uint32_t)w * (uint32_t)v) : v;
              if (first) ai = tap;
              else switch (d->reduce) {
                case HIPACC_REDUCE_MIN: ai = tap < ai ? tap : ai; break;
                case HIPACC_REDUCE_MAX: ai = tap > ai ? tap : ai; break;
                case HIPACC_REDUCE_PROD: ai = (int32_t)((uint32_t)ai * (uint32_t)tap); break;
                default: ai = (int32_t)((uint32_t)ai + (uint32_t)tap); break;
              }
            }
            first = 0;
          }
        }
        const long long o = (long long)(y + out->offset_y) * out->stride + x + out->offset_x;
        if (accf) px_put_f(op, ob, o * nch + ch, af, d->cvt, d->post_div);
        else px_put_i(op, ob, o * nch + ch, ai, d->cvt, d->post_div);
      }
    }
  }
  return 0;
}

/* FP64 accuracy oracle of a float weighted SUM: same taps, double arithmetic,
 * result stored as double (out_f64 dense, out->width x out->height per image). */
ORC_EXPORT int oracle_local_f64(const hipacc_local_desc* d, double* out_f64) {
  const hipacc_view* in = &d->in;
  const hipacc_view* out = &d->out;
  const int mw = d->mask_w, mh = d->mask_h, rx = mw / 2, ry = mh / 2;
  const double cf = const_f(in->dtype, d->const_val);
  for (int b = 0; b < nbatch(out); ++b) {
    const char* ip = plane(in, b);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < out->height; ++y)
      for (int x = 0; x < out->width; ++x) {
        double acc = 0.0;
        for (int dy = 0; dy < mh; ++dy)
          for (int dx = 0; dx < mw; ++dx) {
            const int t = dy * mw + dx;
            if (d->domain && !d->domain[t]) continue;
            long long idx = 0;
            const int hit = bh_tap(in, d->boundary, x + in->offset_x + dx - rx,
                                   y + in->offset_y + dy - ry, &idx);
            const double v = hit ? px_get(ip, in->dtype, idx) : cf;
            double w = 1.0;
            if (d->use_weights) w = d->weights_f32 ? d->weights_f32[t] : (double)d->weights_i32[t];
            acc += w * v;
          }
        out_f64[((long long)b * out->height + y) * out->width + x] = acc;
      }
  }
  return 0;
}

/* ---- separable: the two-kernel DSL program (row pass into an
 * Image<float>/Image<int> the size of the accessor window, then column pass
 * over it with the same boundary mode). ----------------------------------- */
ORC_EXPORT int oracle_separable(const hipacc_separable_desc* d) {
  const hipacc_view* in = &d->in;
  const hipacc_view* out = &d->out;
  const int W = in->width, H = in->height, rx = d->ksize_x / 2, ry = d->ksize_y / 2;
  const int accf = d->acc_dtype == HIPACC_F32;
  const float cf = const_f(in->dtype, d->const_val);
  const int32_t ci = const_i(in->dtype, d->const_val);
  float* tf = accf ? (float*)malloc(sizeof(float) * (size_t)W * H) : NULL;
  int32_t* ti = accf ? NULL : (int32_t*)malloc(sizeof(int32_t) * (size_t)W * H);
  hipacc_view tv;
  memset(&tv, 0, sizeof(tv));
  tv.dtype = accf ? HIPACC_F32 : HIPACC_I32;
  tv.stride = W; tv.img_width = W; tv.img_height = H; tv.width = W; tv.height = H;
  for (int b = 0; b < nbatch(out); ++b) {
    const char* ip = plane(in, b);
    char* op = (char*)plane(out, b);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; ++y)
      for (int x = 0; x < W; ++x) {
        float hf = 0.0f;
        int32_t hi = 0;
        for (int dx = 0; dx < d->ksize_x; ++dx) {
          long long idx = 0;
          const int hit = bh_tap(in, d->boundary, x + in->offset_x + dx - rx, y + in->offset_y, &idx);
          if (accf) {
            const float v = hit ? px_f(ip, in->dtype, idx) : cf;
            if (dx == 0) hf = d->row_f32[dx] * v;
            else if (d->fp_contract) hf = fmaf(d->row_f32[dx], v, hf);
            else { const float p = d->row_f32[dx] * v; hf = hf + p; }
          } else {
            const int32_t v = hit ? px_i(ip, in->dtype, idx) : ci;
            hi = (int32_t)((uint32_t)hi + (uint32_t)d->row_i32[dx] * (uint32_t)v);
          }
        }
        if (accf) tf[(long long)y * W + x] = hf; else ti[(long long)y * W + x] = hi;
      }
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; ++y)
      for (int x = 0; x < W; ++x) {
        float af = 0.0f;
        int32_t ai = 0;
        for (int dy = 0; dy < d->ksize_y; ++dy) {
          long long idx = 0;
          const int hit = bh_tap(&tv, d->boundary, x, y + dy - ry, &idx);
          if (accf) {
            const float v = hit ? tf[idx] : cf;
            if (dy == 0) af = d->col_f32[dy] * v;
            else if (d->fp_contract) af = fmaf(d->col_f32[dy], v, af);
            else { const float p = d->col_f32[dy] * v; af = af + p; }
          } else {
            const int32_t v = hit ? ti[idx] : ci;
            ai = (int32_t)((uint32_t)ai + (uint32_t)d->col_i32[dy] * (uint32_t)v);
          }
        }
        const long long o = (long long)(y + out->offset_y) * out->stride + x + out->offset_x;
        if (accf) px_put_f(op, out->dtype, o, af, d->cvt, d->post_div);
        else px_put_i(op, out->dtype, o, ai, d->cvt, d->post_div);
      }
  }
  free(tf);
  free(ti);
  return 0;
}

/* ---- point operator (Kernel::execute, dsl/kernel.hpp:94-119) ----------- */
ORC_EXPORT int oracle_point(const hipacc_point_desc* d) {
  const hipacc_view* o = &d->out;
  const int binary = d->op >= HIPACC_POINT_ADD && d->op <= HIPACC_POINT_MAGNITUDE;
  const int fmath = d->in0.dtype == HIPACC_F32 || (binary && d->in1.dtype == HIPACC_F32);
  for (int b = 0; b < nbatch(o); ++b) {
    const char* p0 = plane(&d->in0, b);
    const char* p1 = binary ? plane(&d->in1, b) : NULL;
    char* po = (char*)plane(o, b);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < o->height; ++y)
      for (int x = 0; x < o->width; ++x) {
        const long long i0 = (long long)(y + d->in0.offset_y) * d->in0.stride + x + d->in0.offset_x;
        const long long i1 = binary ? (long long)(y + d->in1.offset_y) * d->in1.stride + x + d->in1.offset_x : 0;
        const long long io = (long long)(y + o->offset_y) * o->stride + x + o->offset_x;
        if (fmath) {
          const float a = px_f(p0, d->in0.dtype, i0);
          const float c = binary ? px_f(p1, d->in1.dtype, i1) : 0.0f;
          float r;
          switch (d->op) {
            case HIPACC_POINT_AFFINE: { const float m = d->alpha * a; r = m + d->beta; break; }
            case HIPACC_POINT_ADD: r = a + c; break;
            case HIPACC_POINT_SUB: r = a - c; break;
            case HIPACC_POINT_MUL: r = a * c; break;
            case HIPACC_POINT_ABSDIFF: r = fabsf(a - c); break;
            case HIPACC_POINT_MIN: r = a < c ? a : c; break;
            case HIPACC_POINT_MAX: r = a > c ? a : c; break;
            case HIPACC_POINT_MAGNITUDE: {
              const float aa = a * a, cc = c * c;
              r = sqrtf(aa + cc);
              r = r < 255.0f ? r : 255.0f;
              r = r > 0.0f ? r : 0.0f;
              break;
            }
            default: r = a; break;
          }
          px_put_f(po, o->dtype, io, r, d->cvt, 0.0f);
        } else {
          const int32_t a = px_i(p0, d->in0.dtype, i0);
          const int32_t c = binary ? px_i(p1, d->in1.dtype, i1) : 0;
          if (d->op == HIPACC_POINT_AFFINE) {
            const float m = d->alpha * (float)a;
            px_put_f(po, o->dtype, io, m + d->beta, d->cvt, 0.0f);
          } else if (d->op == HIPACC_POINT_MAGNITUDE) {
            /* Sobel combine: samples-public/3_Preprocessing/Sobel/src/main.cpp:87-96 */
            const int32_t n = (int32_t)d->alpha;
            const int32_t q0 = n ? a / n : a, q1 = n ? c / n : c;
            float r = sqrtf((float)(q0 * q0 + q1 * q1));
            r = r < 255.0f ? r : 255.0f;
            r = r > 0.0f ? r : 0.0f;
            px_put_f(po, o->dtype, io, r, d->cvt, 0.0f);
          } else {
            int32_t r;
            switch (d->op) {
              case HIPACC_POINT_ADD: r = (int32_t)((uint32_t)a + (uint32_t)c); break;
              case HIPACC_POINT_SUB: r = (int32_t)((uint32_t)a - (uint32_t)c); break;
              case HIPACC_POINT_MUL: r = (int32_t)((uint32_t)a * (uint32_t)c); break;
              case HIPACC_POINT_ABSDIFF: r = a > c ? a - c : c - a; break;
              case HIPACC_POINT_MIN: r = a < c ? a : c; break;
              case HIPACC_POINT_MAX: r = a > c ? a : c; break;
              default: r = a; break;
            }
            px_put_i(po, o->dtype, io, r, d->cvt, 0.0f);
          }
        }
      }
  }
  return 0;
}

/* ---- global reduction: REDUCTION_CPU_2D (runtime/hipacc_cpu_red.hpp:19-68)
 * with PPT 1 and `nthreads` partials: each thread's first row seeds its
 * partial with that row's first pixel, partials are combined in thread order.
 * nthreads == 1 is the DSL's sequential order (dsl/kernel.hpp:121-151).
 * Threads that receive no row are skipped in the final combine (the macro
 * folds their zero-initialised partial in, which is only harmless for SUM). */
#define ORC_RED(A, GET)                                                          \
  do {                                                                          \
    A* part = (A*)calloc((size_t)nthreads, sizeof(A));                          \
    int* used = (int*)calloc((size_t)nthreads, sizeof(int));                    \
    _Pragma("omp parallel num_threads(nthreads)")                               \
    {                                                                           \
      int tid = 0;                                                              \
      ORC_TID(tid);                                                             \
      int seeded = 0;                                                           \
      A acc = 0;                                                                \
      _Pragma("omp for schedule(static)")                                       \
      for (int y = 0; y < v->height; ++y) {                                     \
        const long long row = (long long)(y + v->offset_y) * v->stride + v->offset_x; \
        int x0 = 0;                                                             \
        if (!seeded) { acc = (A)GET(v->data, v->dtype, row); seeded = 1; x0 = 1; } \
        for (int x = x0; x < v->width; ++x) {                                   \
          const A p = (A)GET(v->data, v->dtype, row + x);                       \
          switch (mode) {                                                       \
            case HIPACC_REDUCE_MIN: acc = p < acc ? p : acc; break;             \
            case HIPACC_REDUCE_MAX: acc = p > acc ? p : acc; break;             \
            case HIPACC_REDUCE_PROD: acc = acc * p; break;                      \
            default: acc = acc + p; break;                                      \
          }                                                                     \
        }                                                                       \
      }                                                                         \
      part[tid] = acc;                                                          \
      used[tid] = seeded;                                                       \
    }                                                                           \
    int have = 0;                                                               \
    A r = 0;                                                                    \
    for (int t = 0; t < nthreads; ++t) {                                        \
      if (!used[t]) continue;                                                   \
      if (!have) { r = part[t]; have = 1; continue; }                           \
      switch (mode) {                                                           \
        case HIPACC_REDUCE_MIN: r = part[t] < r ? part[t] : r; break;           \
        case HIPACC_REDUCE_MAX: r = part[t] > r ? part[t] : r; break;           \
        case HIPACC_REDUCE_PROD: r = r * part[t]; break;                        \
        default: r = r + part[t]; break;                                        \
      }                                                                         \
    }                                                                           \
    memcpy(result, &r, sizeof(A));                                              \
    free(part);                                                                 \
    free(used);                                                                 \
  } while (0)

#ifdef _OPENMP
#define ORC_TID(t) (t) = omp_get_thread_num()
#else
#define ORC_TID(t) (t) = 0
#endif

static uint32_t px_u(const void* p, int dt, long long i) { return (uint32_t)px_i(p, dt, i); }

ORC_EXPORT int oracle_reduce(const hipacc_view* v, int mode, int acc_dtype, void* result, int nthreads) {
  if (nthreads < 1) nthreads = 1;
  if (acc_dtype == HIPACC_F32) ORC_RED(float, px_f);
  else if (acc_dtype == HIPACC_U32) ORC_RED(uint32_t, px_u);
  else ORC_RED(int32_t, px_i);
  return 0;
}

/* FP64 sum of the window (accuracy reference for float SUM). */
ORC_EXPORT double oracle_sum_f64(const hipacc_view* v) {
  double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
  for (int y = 0; y < v->height; ++y) {
    const long long row = (long long)(y + v->offset_y) * v->stride + v->offset_x;
    for (int x = 0; x < v->width; ++x) s += px_get(v->data, v->dtype, row + x);
  }
  return s;
}

/* ---- binning: BINNING_CPU_2D (runtime/hipacc_cpu_red.hpp:70-126) with the
 * counting binning of the Histogram sample; out-of-range indices dropped
 * as in binPut (:71-76). ------------------------------------------------- */
ORC_EXPORT int oracle_histogram(const hipacc_view* v, int nbins, float scale, uint32_t* bins) {
  int nt = 1;
#ifdef _OPENMP
  nt = omp_get_max_threads();
#endif
  uint32_t* lb = (uint32_t*)calloc((size_t)nt * nbins, sizeof(uint32_t));
#pragma omp parallel num_threads(nt)
  {
    int tid = 0;
    ORC_TID(tid);
    uint32_t* mine = lb + (size_t)tid * nbins;
#pragma omp for schedule(static)
    for (int y = 0; y < v->height; ++y) {
      const long long row = (long long)(y + v->offset_y) * v->stride + v->offset_x;
      for (int x = 0; x < v->width; ++x) {
        long long idx;
        if (v->dtype == HIPACC_U8) idx = ((long long)((const uint8_t*)v->data)[row + x] * nbins) >> 8;
        else if (v->dtype == HIPACC_U16) idx = ((long long)((const uint16_t*)v->data)[row + x] * nbins) >> 16;
        else {
          const float f = ((const float*)v->data)[row + x] * scale;
          if (!(f >= 0.0f) || f >= (float)nbins) continue;
          idx = (long long)(uint32_t)f;
        }
        if (idx >= 0 && idx < nbins) mine[idx] += 1u;
      }
    }
  }
  for (int i = 0; i < nbins; ++i) {
    uint32_t s = 0;
    for (int t = 0; t < nt; ++t) s += lb[(size_t)t * nbins + i];
    bins[i] = s;
  }
  free(lb);
  return 0;
}

/* ---- bilateral: samples-public/3_Preprocessing/Bilateral_Filter/src/
 * main.cpp:63-77, iterate over the full ksize x ksize domain. -------------- */
static void spatial_table(const hipacc_bilateral_desc* d, float* sp) {
  const int r = d->ksize / 2;
  for (int dy = 0; dy < d->ksize; ++dy)
    for (int dx = 0; dx < d->ksize; ++dx) {
      const int t = dy * d->ksize + dx;
      if (d->spatial) sp[t] = d->spatial[t];
      else {
        const double ddx = dx - r, ddy = dy - r;
        sp[t] = (float)exp(-(ddx * ddx + ddy * ddy) / (2.0 * (double)d->sigma_d * d->sigma_d));
      }
    }
}

ORC_EXPORT int oracle_bilateral(const hipacc_bilateral_desc* d) {
  const hipacc_view* in = &d->in;
  const hipacc_view* out = &d->out;
  const int k = d->ksize, r = k / 2;
  float sp[HIPACC_MAX_MASK_DIM * HIPACC_MAX_MASK_DIM];
  spatial_table(d, sp);
  const float c_r = 0.5f / (d->sigma_r * d->sigma_r);
  const float cf = const_f(in->dtype, d->const_val);
  for (int b = 0; b < nbatch(out); ++b) {
    const char* ip = plane(in, b);
    char* op = (char*)plane(out, b);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < out->height; ++y)
      for (int x = 0; x < out->width; ++x) {
        const int cx = x + in->offset_x, cy = y + in->offset_y;
        long long cidx = 0;
        bh_tap(in, HIPACC_BOUNDARY_CLAMP, cx, cy, &cidx);
        const float center = px_f(ip, in->dtype, cidx);
        float dsum = 0.0f, psum = 0.0f;
        for (int dy = 0; dy < k; ++dy)
          for (int dx = 0; dx < k; ++dx) {
            long long idx = 0;
            const int hit = bh_tap(in, d->boundary, cx + dx - r, cy + dy - r, &idx);
            const float v = hit ? px_f(ip, in->dtype, idx) : cf;
            const float diff = v - center;
            const float e = expf(-c_r * diff * diff);
            const float s = e * sp[dy * k + dx];
            dsum += s;
            const float sv = s * v;
            psum += sv;
          }
        const long long o = (long long)(y + out->offset_y) * out->stride + x + out->offset_x;
        px_put_f(op, out->dtype, o, psum / dsum, d->cvt, 0.0f);
      }
  }
  return 0;
}

/* FP64 bilateral (accuracy reference), dense double output. */
ORC_EXPORT int oracle_bilateral_f64(const hipacc_bilateral_desc* d, double* out_f64) {
  const hipacc_view* in = &d->in;
  const hipacc_view* out = &d->out;
  const int k = d->ksize, r = k / 2;
  float sp[HIPACC_MAX_MASK_DIM * HIPACC_MAX_MASK_DIM];
  spatial_table(d, sp);
  const double c_r = 0.5 / ((double)d->sigma_r * d->sigma_r);
  const double cf = const_f(in->dtype, d->const_val);
  for (int b = 0; b < nbatch(out); ++b) {
    const char* ip = plane(in, b);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < out->height; ++y)
      for (int x = 0; x < out->width; ++x) {
        const int cx = x + in->offset_x, cy = y + in->offset_y;
        long long cidx = 0;
        bh_tap(in, HIPACC_BOUNDARY_CLAMP, cx, cy, &cidx);
        const double center = px_get(ip, in->dtype, cidx);
        double dsum = 0.0, psum = 0.0;
        for (int dy = 0; dy < k; ++dy)
          for (int dx = 0; dx < k; ++dx) {
            long long idx = 0;
            const int hit = bh_tap(in, d->boundary, cx + dx - r, cy + dy - r, &idx);
            const double v = hit ? px_get(ip, in->dtype, idx) : cf;
            const double s = exp(-c_r * (v - center) * (v - center)) * (double)sp[dy * k + dx];
            dsum += s;
            psum += s * v;
          }
        out_f64[((long long)b * out->height + y) * out->width + x] = psum / dsum;
      }
  }
  return 0;
}

/* ---- resampling: Interpolation::interpolate (dsl/image.hpp:390-422) for
 * Interpolate::NN / LF, in float with the DSL's evaluation order (this file
 * is compiled without contraction), reading through the accessor's boundary
 * handling; then COPY / SUB (in2 - v) / ADD (v + in2). -------------------- */
/* Interpolation weights of dsl/image.hpp:316-382 (binomial_5,
 * bicubic_spline with a = -0.5, lanczos with l = 3, the last evaluated in
 * double and rounded once to float), in the source's operation order.      */
static float w_b5(float d) {
  d = fabsf(d);
  if (d < 0.5f) return 6.0f / 8.0f;
  if (d < 1.0f) return 4.0f / 8.0f;
  if (d < 1.5f) return 1.0f / 8.0f;
  return 0.0f;
}
static float w_cf(float d) {
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
static float w_l3(float d) {
  const float l = 3.0f;
  d = fabsf(d);
  if (d == 0.0f) return 1.0f;
  if (d < l) {
    const double pi = atan(1.0) * 4;
    const double num = (double)l * (sin(pi * d / l) * sin(pi * d));
    const double den = pi * pi * d * d;
    return (float)(num / den);
  }
  return 0.0f;
}

/* Interpolation::interpolate (dsl/image.hpp:390-532) for the modes NN, LF,
 * B5, CF, L3 of an accessor window read with its boundary condition; the
 * iteration-space pixel x is read at x + xf (the accessor call in(xf, yf)).
 * Sums run left to right with a rounding per product and per add: rows of
 * taps first (y0 .. y3 / y5), then the row sums weighted by the y weights.
 * L3 keeps the reference's row weights lanczos(y_frac - 2 + k) for rows
 * y_int - 1 + k and its second row's fifth tap weight lanczos(x_frac - 2 + 5)
 * (dsl/image.hpp:493; the generated CUDA code has the same,
 * runtime/hipacc_cu_interpolate.hpp:491).                                   */
static float rs_px(const hipacc_resample_desc* d, int x, int y, float cf) {
  long long idx = 0;
  return bh_tap(&d->in, d->boundary, x, y, &idx) ? px_f(d->in.data, d->in.dtype, idx) : cf;
}
ORC_EXPORT int oracle_resample(const hipacc_resample_desc* d) {
  const hipacc_view* in = &d->in;
  const hipacc_view* out = &d->out;
  const float stride_x = in->width / (float)out->width;
  const float stride_y = in->height / (float)out->height;
  const float cf = const_f(in->dtype, d->const_val);
  for (int y = 0; y < out->height; ++y)
    for (int x = 0; x < out->width; ++x) {
      const float bx = in->offset_x + stride_x / 2, by = in->offset_y + stride_y / 2;
      const float x_mapped = bx + stride_x * (float)(x + d->xf);
      const float y_mapped = by + stride_y * (float)(y + d->yf);
      float xb = x_mapped - 0.5f, yb = y_mapped - 0.5f;
      if (xb < 0.0f) xb = 0.0f;
      if (yb < 0.0f) yb = 0.0f;
      const int xi = (int)xb, yi = (int)yb;
      float xf = xb - xi, yf = yb - yi;
      float v = 0.0f;
      switch (d->interp) {
        case HIPACC_INTERP_NN:
          v = rs_px(d, (int)x_mapped, (int)y_mapped, cf);
          break;
        case HIPACC_INTERP_LF: {
          const float p00 = rs_px(d, xi, yi, cf), p10 = rs_px(d, xi + 1, yi, cf);
          const float p01 = rs_px(d, xi, yi + 1, cf), p11 = rs_px(d, xi + 1, yi + 1, cf);
          v = (1.0f - xf) * (1.0f - yf) * p00 + xf * (1.0f - yf) * p10 +
              (1.0f - xf) * yf * p01 + xf * yf * p11;
          break;
        }
        case HIPACC_INTERP_B5: {
          xf += 0.5f;
          yf += 0.5f;
          float r[4];
          for (int k = 0; k < 4; ++k) {
            float a = rs_px(d, xi, yi + k, cf) * w_b5(xf - 0);
            for (int j = 1; j < 4; ++j) a = a + rs_px(d, xi + j, yi + k, cf) * w_b5(xf - (float)j);
            r[k] = a;
          }
          v = r[0] * w_b5(yf - 0);
          for (int k = 1; k < 4; ++k) v = v + r[k] * w_b5(yf - (float)k);
          break;
        }
        case HIPACC_INTERP_CF: {
          float r[4];
          for (int k = 0; k < 4; ++k) {
            float a = rs_px(d, xi - 1, yi - 1 + k, cf) * w_cf((xf - 1.0f) + 0.0f);
            for (int j = 1; j < 4; ++j)
              a = a + rs_px(d, xi - 1 + j, yi - 1 + k, cf) * w_cf((xf - 1.0f) + (float)j);
            r[k] = a;
          }
          v = r[0] * w_cf((yf - 1.0f) + 0.0f);
          for (int k = 1; k < 4; ++k) v = v + r[k] * w_cf((yf - 1.0f) + (float)k);
          break;
        }
        case HIPACC_INTERP_L3: {
          float r[6];
          for (int k = 0; k < 6; ++k) {
            float a = rs_px(d, xi - 2, yi - 1 + k, cf) * w_l3((xf - 2.0f) + 0.0f);
            for (int j = 1; j < 6; ++j) {
              const int jw = (k == 1 && j == 4) ? 5 : j;
              a = a + rs_px(d, xi - 2 + j, yi - 1 + k, cf) * w_l3((xf - 2.0f) + (float)jw);
            }
            r[k] = a;
          }
          v = r[0] * w_l3((yf - 2.0f) + 0.0f);
          for (int k = 1; k < 6; ++k) v = v + r[k] * w_l3((yf - 2.0f) + (float)k);
          break;
        }
        default: return 1;
      }
      if (d->op != HIPACC_RESAMPLE_COPY) {
        const float b = px_f(d->in2.data, d->in2.dtype,
                             (long long)(y + d->in2.offset_y) * d->in2.stride + x + d->in2.offset_x);
        v = d->op == HIPACC_RESAMPLE_SUB ? b - v : v + b;
      }
      px_put_f(out->data, out->dtype, (long long)(y + out->offset_y) * out->stride + x + out->offset_x,
               v, HIPACC_CVT_TRUNC, 0.0f);
    }
  return 0;
}

/* ---- adjoint of the float weighted SUM (no reference counterpart: the
 * reference has no gradients).  Scatter form, the literal transpose of the
 * forward's index map (dsl/kernel.hpp:241-267 with the boundary maps of
 * lib/AST/BorderHandling.cpp:41-120): every tap of every output pixel sends
 * w[t]*dY(p) to the input pixel it read and dY(p)*X to dw[t].  In double;
 * dx is dense [batch][h][w].  dx_abs / dw_abs (optional) receive the sums of
 * the absolute values of the terms, the scale of each result's rounding. */
ORC_EXPORT int oracle_local_grad(const hipacc_local_grad_desc* d, double* dx, double* dw,
                                 double* dx_abs, double* dw_abs) {
  const hipacc_view* in = &d->in;
  const hipacc_view* gy = &d->grad_out;
  const int W = in->width, H = in->height, kx = d->mask_w, ky = d->mask_h;
  const int rx = kx / 2, ry = ky / 2, nb = nbatch(in);
  const float* w = d->weights_f32;
  const double cval = (double)(float)d->const_val;
  memset(dx, 0, sizeof(double) * (size_t)nb * W * H);
  memset(dw, 0, sizeof(double) * kx * ky);
  if (dx_abs) memset(dx_abs, 0, sizeof(double) * (size_t)nb * W * H);
  if (dw_abs) memset(dw_abs, 0, sizeof(double) * kx * ky);
  for (int b = 0; b < nb; ++b) {
    const float* X = (const float*)plane(in, b);
    const float* G = (const float*)plane(gy, b);
    double* DX = dx + (size_t)b * W * H;
    double* DXA = dx_abs ? dx_abs + (size_t)b * W * H : NULL;
    for (int y = 0; y < H; ++y)
      for (int x = 0; x < W; ++x) {
        const double g = G[(long long)y * gy->stride + x];
        for (int ty = 0; ty < ky; ++ty)
          for (int tx = 0; tx < kx; ++tx) {
            const int t = ty * kx + tx;
            const int sx = x + tx - rx, sy = y + ty - ry;
            if (d->boundary == HIPACC_BOUNDARY_CONSTANT &&
                (sx < 0 || sx >= W || sy < 0 || sy >= H)) {
              dw[t] += g * cval;
              if (dw_abs) dw_abs[t] += fabs(g * cval);
              continue;
            }
            const int ix = bh_index(d->boundary, sx, 0, W, W);
            const int iy = bh_index(d->boundary, sy, 0, H, H);
            const double xv = X[(long long)iy * in->stride + ix];
            DX[(long long)iy * W + ix] += (double)w[t] * g;
            dw[t] += g * xv;
            if (DXA) DXA[(long long)iy * W + ix] += fabs((double)w[t] * g);
            if (dw_abs) dw_abs[t] += fabs(g * xv);
          }
      }
  }
  return 0;
}

/* ---- bilateral backward (no reference counterpart).  For each output pixel
 * p the forward of oracle_bilateral_f64 is re-evaluated in double and the
 * chain rule applied tap by tap in the forward's order: the tap value v read
 * at X(q) receives a*s*(1 - 2c (v - X(p)) (v - Y(p))), the centre X(p)
 * receives 2c a SUM s (v - X(p)) (v - Y(p)), and dL/dc collects
 * -a SUM s (v - X(p))^2 (v - Y(p)), with a = dY(p)/D(p), c = 0.5/sigma_r^2.
 * Outputs: dx dense [batch][h][w], *dsigma = dL/dc * (-1/sigma_r^3); the
 * optional dx_abs and *dsigma_abs are the sums of the absolute values of the
 * terms (the scale of each result's rounding). */
ORC_EXPORT int oracle_bilateral_grad(const hipacc_bilateral_grad_desc* d, double* dx,
                                     double* dsigma, double* dx_abs, double* dsigma_abs) {
  const hipacc_view* in = &d->in;
  const hipacc_view* gy = &d->grad_out;
  const int W = in->width, H = in->height, k = d->ksize, r = k / 2, nb = nbatch(in);
  hipacc_bilateral_desc bd;
  memset(&bd, 0, sizeof(bd));
  bd.ksize = d->ksize;
  bd.sigma_d = d->sigma_d;
  bd.spatial = d->spatial;
  float sp[HIPACC_MAX_MASK_DIM * HIPACC_MAX_MASK_DIM];
  spatial_table(&bd, sp);
  const double c = 0.5 / ((double)d->sigma_r * d->sigma_r);
  const double cv = (double)(float)d->const_val;
  double dc = 0.0, dc_abs = 0.0;
  memset(dx, 0, sizeof(double) * (size_t)nb * W * H);
  if (dx_abs) memset(dx_abs, 0, sizeof(double) * (size_t)nb * W * H);
  for (int b = 0; b < nb; ++b) {
    const float* X = (const float*)plane(in, b);
    const float* G = (const float*)plane(gy, b);
    double* DX = dx + (size_t)b * W * H;
    double* DXA = dx_abs ? dx_abs + (size_t)b * W * H : NULL;
    for (int y = 0; y < H; ++y)
      for (int x = 0; x < W; ++x) {
        const double xp = X[(long long)y * in->stride + x];
        double v[HIPACC_MAX_MASK_DIM * HIPACC_MAX_MASK_DIM], s[HIPACC_MAX_MASK_DIM * HIPACC_MAX_MASK_DIM];
        long long q[HIPACC_MAX_MASK_DIM * HIPACC_MAX_MASK_DIM];
        double D = 0.0, N = 0.0;
        for (int ty = 0; ty < k; ++ty)
          for (int tx = 0; tx < k; ++tx) {
            const int t = ty * k + tx;
            const int sx = x + tx - r, sy = y + ty - r;
            if (d->boundary == HIPACC_BOUNDARY_CONSTANT && (sx < 0 || sx >= W || sy < 0 || sy >= H)) {
              q[t] = -1;
              v[t] = cv;
            } else {
              const int ix = bh_index(d->boundary, sx, 0, W, W);
              const int iy = bh_index(d->boundary, sy, 0, H, H);
              q[t] = (long long)iy * W + ix;
              v[t] = X[(long long)iy * in->stride + ix];
            }
            s[t] = exp(-c * (v[t] - xp) * (v[t] - xp)) * (double)sp[t];
            D += s[t];
            N += s[t] * v[t];
          }
        const double Y = N / D;
        const double a = (double)G[(long long)y * gy->stride + x] / D;
        double centre = 0.0, centre_abs = 0.0;
        for (int t = 0; t < k * k; ++t) {
          const double dv = v[t] - xp, ey = v[t] - Y;
          if (q[t] >= 0) {
            const double term = a * s[t] * (1.0 - 2.0 * c * dv * ey);
            DX[q[t]] += term;
            if (DXA) DXA[q[t]] += fabs(term);
          }
          centre += s[t] * dv * ey;
          centre_abs += fabs(s[t] * dv * ey);
          dc -= a * s[t] * dv * dv * ey;
          dc_abs += fabs(a * s[t] * dv * dv * ey);
        }
        DX[(long long)y * W + x] += 2.0 * c * a * centre;
        if (DXA) DXA[(long long)y * W + x] += fabs(2.0 * c * a) * centre_abs;
      }
  }
  const double sr = d->sigma_r;
  *dsigma = dc * (-1.0 / (sr * sr * sr));
  if (dsigma_abs) *dsigma_abs = dc_abs / (sr * sr * sr);
  return 0;
}

/* ---- synthetic pixels (same generator as the GPU fill kernel) ---------- */
ORC_EXPORT int oracle_fill_random(const hipacc_view* v, uint64_t seed) {
  for (int b = 0; b < nbatch(v); ++b) {
    char* p = (char*)plane(v, b);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < v->height; ++y)
      for (int x = 0; x < v->width; ++x) {
        const uint64_t k = (uint64_t)b * v->width * v->height + (uint64_t)y * v->width + x;
        const uint64_t r = sm64_at(seed, k + 1);
        const long long i = (long long)(y + v->offset_y) * v->stride + x + v->offset_x;
        switch (v->dtype) {
          case HIPACC_U8: ((uint8_t*)p)[i] = (uint8_t)(r >> 56); break;
          case HIPACC_I8: ((int8_t*)p)[i] = (int8_t)(uint8_t)(r >> 56); break;
          case HIPACC_U16: ((uint16_t*)p)[i] = (uint16_t)(r >> 48); break;
          case HIPACC_I16: ((int16_t*)p)[i] = (int16_t)(uint16_t)(r >> 48); break;
          case HIPACC_U32: ((uint32_t*)p)[i] = (uint32_t)(r >> 32); break;
          case HIPACC_I32: ((int32_t*)p)[i] = (int32_t)(uint32_t)(r >> 32); break;
          default: ((float*)p)[i] = (float)(r >> 40) * (1.0f / 16777216.0f); break;
        }
      }
  }
  return 0;
}

ORC_EXPORT int oracle_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
