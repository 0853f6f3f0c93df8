/*
 * hipacc_hip.h -- C ABI of libhipacc_gfx950.so, the MI355X (gfx950) execution
 * library for Hipacc's data-parallel operators.
 *
 * The reference executes DSL operators through generated kernels launched by
 * its CUDA runtime (runtime/hipacc_cu.hpp, runtime/hipacc_cu.tpp,
 * runtime/hipacc_cu_standalone.hpp, runtime/hipacc_cu_red.hpp).  This library
 * replaces the generated-kernel layer: instead of one generated kernel per DSL
 * operator, the host side describes the operator (accessor window, iteration
 * space, mask/domain, reduction mode, boundary mode, output conversion) and
 * calls one of the entry points below, which launch hand-written gfx950
 * kernels.
 *
 * Conventions
 *  - Every entry point returns an int status (HIPACC_OK == 0) and never
 *    aborts; hipacc_hip_strerror() names the code, hipacc_hip_last_error()
 *    gives the detailed message of the calling thread's last failure.
 *  - Pixel pointers are device pointers; mask / domain tables are host
 *    pointers (copied into the kernel arguments at launch), except the
 *    explicitly named *_dev members.
 *  - `stream` is a hipStream_t passed as void* (NULL = the null stream).
 *    Launches are asynchronous on that stream; nothing here synchronises
 *    unless its comment says so.
 *  - Calls on different streams may run concurrently.  Histograms use a
 *    workspace per (device, stream), so they are reentrant per stream.
 *    hipacc_hip_reduce_sync holds its device's lock until its result is on
 *    the host.  The other users of a per-device library workspace
 *    (hipacc_hip_reduce and the adjoints with a NULL workspace, and
 *    hipacc_hip_local_grad_noalloc) take the lock at enqueue time only:
 *    concurrent calls of those on different streams must pass their own
 *    workspace (the no-allocation form has none and must not overlap).
 */
#ifndef HIPACC_HIP_H
#define HIPACC_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
enum {
  HIPACC_OK = 0,
  HIPACC_ERR_INVALID_ARG = 1,   /* a descriptor field is out of range      */
  HIPACC_ERR_UNSUPPORTED = 2,   /* valid, but not implemented by a kernel  */
  HIPACC_ERR_HIP = 3,           /* a HIP runtime call failed               */
  HIPACC_ERR_NO_DEVICE = 4,     /* no gfx950 device / init failed          */
  HIPACC_ERR_OUT_OF_MEMORY = 5
};

/* ---- pixel types ------------------------------------------------------- */
/* The scalar pixel types of the DSL (dsl/types.hpp:33-36 uchar/ushort/uint
 * plus char/short/int/float), and its 4-wide vector pixels (dsl/types.hpp:
 * 38-100: uchar4, short4, int4, float4; four interleaved channels per
 * pixel, `stride` and offsets counted in pixels).  A vector operator applies
 * the scalar operator to each channel: the DSL's vector arithmetic is
 * componentwise.                                                           */
typedef enum {
  HIPACC_U8 = 0, HIPACC_I8 = 1, HIPACC_U16 = 2, HIPACC_I16 = 3,
  HIPACC_U32 = 4, HIPACC_I32 = 5, HIPACC_F32 = 6,
  HIPACC_U8X4 = 7, HIPACC_I16X4 = 8, HIPACC_I32X4 = 9, HIPACC_F32X4 = 10
} hipacc_dtype;

/* ---- boundary modes: same values as hipacc::Boundary (dsl/image.hpp:46-52) */
typedef enum {
  HIPACC_BOUNDARY_UNDEFINED = 0,
  HIPACC_BOUNDARY_CLAMP = 1,
  HIPACC_BOUNDARY_REPEAT = 2,
  HIPACC_BOUNDARY_MIRROR = 3,
  HIPACC_BOUNDARY_CONSTANT = 4
} hipacc_boundary;

/* ---- reduction modes: same values as hipacc::Reduce (dsl/kernel.hpp:48-54);
 * MEDIAN is unsupported there (dsl/kernel.hpp:259) and here.               */
typedef enum {
  HIPACC_REDUCE_SUM = 0, HIPACC_REDUCE_MIN = 1, HIPACC_REDUCE_MAX = 2,
  HIPACC_REDUCE_PROD = 3
} hipacc_reduce;

/* ---- accumulator -> output conversion --------------------------------- */
/* TRUNC is the C cast the DSL applies on `output() = acc`
 * (dsl/types.hpp:115-117 convert<T>(float) == (T)s).  ROUND is the samples'
 * idiom `(uchar)(acc + 0.5f)` (e.g. samples-public/1_Local_Operators/
 * Gaussian_Blur/src/main.cpp:63-65).  Float-to-integer conversion truncates
 * toward zero; for values the output type can hold that is the C cast.  The
 * reference leaves the rest undefined; here truncation saturates to the
 * uint32 range for uint outputs and to the int32 range for every other
 * integer output (NaN -> 0), which then wraps to the output width.          */
typedef enum { HIPACC_CVT_TRUNC = 0, HIPACC_CVT_ROUND = 1 } hipacc_cvt;

/* ---- an image window: an Accessor or an IterationSpace ----------------- */
/* Mirrors HipaccAccessor / hipacc_launch_info geometry
 * (runtime/hipacc_base.hpp:72-101, runtime/hipacc_cu.hpp:162-174): the
 * window [offset_x, offset_x+width) x [offset_y, offset_y+height) of an image
 * of img_width x img_height pixels with a row pitch of `stride` elements.
 * For an Accessor the window is the region boundary handling refers to
 * (lib/AST/BorderHandling.cpp:155-168); for an IterationSpace it is the set
 * of output pixels (dsl/iterationspace.hpp:45-138).
 * `batch` images are laid out `plane_stride` elements apart (batch 0 or 1 ==
 * a single image).                                                          */
typedef struct hipacc_view {
  void*   data;          /* device pointer to pixel (0,0) of image 0       */
  int32_t dtype;         /* hipacc_dtype                                    */
  int32_t stride;        /* row pitch, elements                             */
  int32_t img_width;     /* allocation extent                               */
  int32_t img_height;
  int32_t width;         /* window                                          */
  int32_t height;
  int32_t offset_x;
  int32_t offset_y;
  int32_t batch;         /* number of images (0 == 1)                       */
  int64_t plane_stride;  /* elements between consecutive images             */
} hipacc_view;

/* ---- local operator: Kernel::convolve / reduce(Domain) / iterate ------ */
/* Computes, for every pixel p of the iteration space `out`,
 *     acc = (+)_{taps t in row-major order, domain[t] != 0} tap(t)
 *     tap(t) = weight[t] * in_bh(p + t)    (use_weights != 0)
 *            = in_bh(p + t)                (use_weights == 0)
 *     out(p) = cvt(post_div != 0 ? (float)acc / post_div : acc)
 * with (+) the reduction `reduce`, seeded by the first active tap as in
 * Kernel::convolve / Kernel::reduce (dsl/kernel.hpp:241-296), taps in the
 * order of ElementIterator / DomainIterator (dsl/iterationspace.hpp:78-90,
 * dsl/mask.hpp:112-126), and in_bh the accessor read with boundary handling
 * (dsl/image.hpp:574-612; generated form lib/AST/BorderHandling.cpp:41-120).
 * The input coordinate of output pixel (x, y) is
 * (x - out.offset_x + in.offset_x, y - out.offset_y + in.offset_y)
 * (dsl/image.hpp:390-412, Interpolate::NO).
 * Vector pixels (in and out both vector typed; host weights): each channel
 * is this operator on that channel, with the constant of CONSTANT in every
 * channel -- the RGBA samples (samples-public/1_Local_Operators/
 * Gaussian_Blur_RGBA, Box_Blur_RGBA, Erode_RGBA, Dilate_RGBA, Laplace_RGBA,
 * 3_Preprocessing/Sobel_RGBA) in the DSL's componentwise arithmetic.
 * acc_dtype is HIPACC_F32 (float lambda, e.g. Gaussian) or HIPACC_I32
 * (int lambda, e.g. Sobel / box sum).  fp_contract = 1 evaluates a float
 * weighted SUM as acc = fma(w, x, acc), the contraction the reference CUDA
 * backend gets from nvcc's default -fmad=true; 0 evaluates acc + w*x with
 * two roundings, as the DSL host emulation compiled by g++ for x86-64 does.
 * Replaces: one generated kernel per convolve() call site
 * (lib/AST/Convolution.cpp:228-467, launched via hipaccLaunchKernel,
 * runtime/hipacc_cu_standalone.hpp:277-329).                               */
#define HIPACC_MAX_MASK_DIM 15
typedef struct hipacc_local_desc {
  hipacc_view in;            /* Accessor                                    */
  hipacc_view out;           /* IterationSpace                              */
  int32_t mask_w, mask_h;    /* odd, 1..HIPACC_MAX_MASK_DIM                 */
  const float*   weights_f32;/* host, mask_h*mask_w, row-major (float mask) */
  const int32_t* weights_i32;/* host, integer mask (acc_dtype I32)          */
  const float*   weights_dev;/* device, float mask (learnable weights);
                                takes precedence over the host tables       */
  const uint8_t* domain;     /* host, nonzero = tap active; NULL = all      */
  int32_t use_weights;       /* 0: tap = pixel (Domain reduce)              */
  int32_t acc_dtype;         /* HIPACC_F32 or HIPACC_I32                    */
  int32_t reduce;            /* hipacc_reduce                               */
  int32_t boundary;          /* hipacc_boundary                             */
  double  const_val;         /* CONSTANT value, in the input pixel type     */
  int32_t cvt;               /* hipacc_cvt                                  */
  float   post_div;          /* 0 = none                                    */
  int32_t fp_contract;       /* see above                                   */
} hipacc_local_desc;
int hipacc_hip_local(const hipacc_local_desc* desc, void* stream);

/* ---- fused operator chains ------------------------------------------- */
/* Kernel fusion (the reference's compiler pass lib/AST/ASTFuse.cpp:212-413
 * with its point-to-point, point-to-local, local-to-point and
 * local-to-local transformations, chains found by lib/Analysis/
 * HostDataDeps.cpp:903-1374; samples-public/6_Test/Kernel_Fusion_{P2P,P2L,
 * L2P,L2L,Mixed}): a chain of same-size kernels, each reading the image
 * the previous one wrote, runs as one launch without the intermediate
 * images, and its result equals the unfused chain's bit for bit.
 *
 * stages[0] reads `in`; stages[k] reads the image stages[k-1] writes;
 * stages[num_stages-1] writes `out`.  All windows (in, out, in1) have the
 * size of `out`; the intermediate images are out-sized images at offset 0.
 *   POINT  hipacc_hip_point's arithmetic: op, alpha, beta, cvt on the
 *          previous image and, for a two-input op, on `in1` (same pixel
 *          type as the previous image), read at the stage's own pixel.
 *   LOCAL  a local operator (`local`: mask, weights, domain, reduce,
 *          acc_dtype, boundary, const_val, cvt, post_div, fp_contract;
 *          its `in` and `out` views are ignored) over the previous image
 *          with that image's boundary handling: a local first stage reads
 *          `in` through its window as hipacc_hip_local does; a later one
 *          reads an intermediate image (UNDEFINED is clamped to it).
 *          At most two LOCAL stages, as in the reference (host weights).
 * Pixel types of every image: U8, I8, U16, I16, I32, F32; out.dtype must
 * be the last stage's dtype; `out` must not overlap `in` or any `in1`.   */
#define HIPACC_MAX_FUSED_STAGES 8
#define HIPACC_MAX_FUSED_LOCALS 2
typedef enum hipacc_stage_kind { HIPACC_STAGE_POINT = 0, HIPACC_STAGE_LOCAL = 1 } hipacc_stage_kind;
typedef struct hipacc_fused_stage {
  int32_t kind;              /* hipacc_stage_kind                             */
  int32_t dtype;             /* pixel type of the image the stage writes      */
  int32_t op;                /* POINT: hipacc_point_op                        */
  float   alpha, beta;       /* POINT                                         */
  int32_t cvt;               /* POINT: hipacc_cvt                             */
  hipacc_view in1;           /* POINT: second input of a two-input op         */
  const hipacc_local_desc* local;   /* LOCAL                                  */
} hipacc_fused_stage;
typedef struct hipacc_fused_desc {
  hipacc_view in;            /* the first stage's accessor                    */
  hipacc_view out;           /* the last stage's iteration space              */
  int32_t num_stages;        /* 1..HIPACC_MAX_FUSED_STAGES                    */
  hipacc_fused_stage stages[HIPACC_MAX_FUSED_STAGES];
} hipacc_fused_desc;
int hipacc_hip_fused(const hipacc_fused_desc* desc, void* stream);

/* ---- separable local operator: a row pass then a column pass ---------- */
/* The two-kernel DSL program
 *     tmp(x, y) = SUM_i row[i] * in_bh(x + i - rx, y)          (float)
 *     out(x, y) = cvt( SUM_j col[j] * tmp_bh(x, y + j - ry) )
 * with tmp an Image<float> the size of the accessor window and both kernels
 * using `boundary` (for CONSTANT, tmp rows outside the window read
 * const_val), fused into one launch that keeps tmp in registers.
 * acc_dtype HIPACC_I32 with integer row/col weights computes the same two
 * passes in 32-bit integers (exact, so equal to the 2D integer mask
 * row (x) col, e.g. a box sum).  The iteration space must have the size of
 * the accessor window.  Replaces two generated convolve() kernels.         */
typedef struct hipacc_separable_desc {
  hipacc_view in;
  hipacc_view out;
  int32_t ksize_x, ksize_y;   /* odd, 1..2*HIPACC_MAX_SEP_RADIUS+1          */
  const float*   row_f32;     /* host, ksize_x weights                      */
  const float*   col_f32;     /* host, ksize_y weights                      */
  const int32_t* row_i32;     /* host, integer weights (acc_dtype I32)      */
  const int32_t* col_i32;
  int32_t acc_dtype;
  int32_t boundary;
  double  const_val;
  int32_t cvt;
  float   post_div;
  int32_t fp_contract;
} hipacc_separable_desc;
#define HIPACC_MAX_SEP_RADIUS 12
int hipacc_hip_separable(const hipacc_separable_desc* desc, void* stream);

/* ---- point operator: Kernel::execute with a per-pixel output() -------- */
/* Per pixel of the iteration space `out` (dsl/kernel.hpp:94-119,
 * dsl/iterationspace.hpp:78-90), reading in0 (and in1) at the same
 * iteration-space-relative coordinate (no boundary handling):
 *   COPY     out = cvt(a)                 AFFINE  out = cvt(alpha*a + beta)
 *   ADD      out = cvt(a + b)             SUB     out = cvt(a - b)
 *   MUL      out = cvt(a * b)             ABSDIFF out = cvt(|a - b|)
 *   MIN/MAX  out = cvt(min/max(a, b))
 *   MAGNITUDE out = cvt(clamp(sqrtf((float)(a'*a' + b'*b')), 0, 255)) with
 *            a' = a / (int)alpha for integer inputs (C division), as the
 *            Sobel combine kernel (samples-public/3_Preprocessing/Sobel/
 *            src/main.cpp:87-96).
 * Arithmetic is int32 when both inputs are integer types, float otherwise,
 * matching C++'s usual arithmetic conversions for these types.
 * Vector pixels (all views vector typed): each channel, as the scalar
 * operator (the Sobel_RGBA combine, samples-public/3_Preprocessing/
 * Sobel_RGBA).                                                              */
typedef enum {
  HIPACC_POINT_COPY = 0, HIPACC_POINT_AFFINE = 1, HIPACC_POINT_ADD = 2,
  HIPACC_POINT_SUB = 3, HIPACC_POINT_MUL = 4, HIPACC_POINT_ABSDIFF = 5,
  HIPACC_POINT_MIN = 6, HIPACC_POINT_MAX = 7, HIPACC_POINT_MAGNITUDE = 8
} hipacc_point_op;
typedef struct hipacc_point_desc {
  hipacc_view in0;
  hipacc_view in1;            /* data == NULL for unary ops                  */
  hipacc_view out;
  int32_t op;
  float   alpha, beta;
  int32_t cvt;
} hipacc_point_desc;
int hipacc_hip_point(const hipacc_point_desc* desc, void* stream);

/* ---- global reduction: Kernel::reduced_data() ------------------------- */
/* Reduces the window of `img` with `reduce` (SUM/MIN/MAX/PROD) and writes
 * one value of type acc_dtype to the device pointer `result_dev`.  SUM of an
 * integer image accumulates in acc_dtype (e.g. U8 image, U32 sum: the DSL
 * program with a Kernel<uint> point stage).  Float SUM is evaluated as a
 * fixed-shape tree (deterministic run to run); the reference's own CPU
 * result depends on the OpenMP thread count (runtime/hipacc_cpu_red.hpp:
 * 19-68), so float sums agree with it only to a tolerance.
 * workspace: device memory of hipacc_hip_reduce_workspace_bytes() bytes, or
 * NULL to use the library's per-device workspace.
 * Replaces hipaccApplyReductionShared (runtime/hipacc_cu.tpp:311-408) and
 * the kernel hipacc_shared_reduction (runtime/hipacc_cu_red.hpp:336-346).  */
size_t hipacc_hip_reduce_workspace_bytes(void);
int hipacc_hip_reduce(const hipacc_view* img, int32_t reduce, int32_t acc_dtype,
                      void* result_dev, void* workspace, void* stream);
/* Blocking convenience form: result copied to host memory `result_host`
 * (acc_dtype-sized), as the reference's reduced_data() returns by value.   */
int hipacc_hip_reduce_sync(const hipacc_view* img, int32_t reduce,
                           int32_t acc_dtype, void* result_host, void* stream);

/* ---- binning: Kernel::binned_data() for a counting histogram ---------- */
/* bin(idx) = 1 with reduce(a, b) = a + b (samples-public/2_Global_Operators/
 * Histogram/src/main.cpp:45-66).  idx = (p * num_bins) >> 8 for U8, >> 16
 * for U16; idx = (uint)(p * scale) for F32 (p < 0 or NaN dropped).  Indices
 * >= num_bins are dropped, as the generated CPU code does
 * (runtime/hipacc_cpu_red.hpp:71-76).  bins_dev (device, num_bins uint32)
 * is overwritten (every bin is written; it need not be cleared).  Two
 * launches: per-block LDS histograms into a workspace of the (device,
 * stream) pair, then one block per bin sums them.
 * Replaces hipaccApplyBinningSegmented
 * (runtime/hipacc_cu.tpp:410-464) and hipacc_binning_reduction
 * (runtime/hipacc_cu_red.hpp:495-641).                                      */
#define HIPACC_MAX_BINS 4096
int hipacc_hip_histogram(const hipacc_view* img, int32_t num_bins, float scale,
                         uint32_t* bins_dev, void* stream);

/* ---- bilateral filter ------------------------------------------------- */
/* The Bilateral_Filter sample's kernel (samples-public/3_Preprocessing/
 * Bilateral_Filter/src/main.cpp:63-77) over a ksize x ksize domain:
 *     c_r = 0.5f / (sigma_r*sigma_r); center = in(p)
 *     s_t = expf(-c_r * (in(p+t) - center)^2) * spatial[t]
 *     out(p) = cvt( (SUM s_t * in(p+t)) / (SUM s_t) )
 * spatial: host ksize*ksize table, or NULL for
 * spatial[t] = (float)exp(-(dx^2+dy^2) / (2*sigma_d^2)) (double exp).
 * F32 in / F32 out, or U8 in / U8 out (cvt ROUND as in the sample).        */
typedef struct hipacc_bilateral_desc {
  hipacc_view in;
  hipacc_view out;
  int32_t ksize;             /* odd, 1..HIPACC_MAX_MASK_DIM                  */
  float   sigma_d;
  float   sigma_r;
  const float* spatial;
  int32_t boundary;
  double  const_val;
  int32_t cvt;
} hipacc_bilateral_desc;
int hipacc_hip_bilateral(const hipacc_bilateral_desc* desc, void* stream);

/* ---- adjoint of a learnable float convolution ------------------------- */
/* The forward is the float weighted SUM of hipacc_hip_local over whole
 * images (accessor window = image, iteration space of the same size):
 *     Y(p) = SUM_t w[t] * X_bh(p + t - r)          (dsl/kernel.hpp:241-267)
 * Given dY = dL/dY this computes
 *     dX(q) = SUM_{p,t : bh(p + t - r) = q} w[t] * dY(p)
 *     dw[t] = SUM_p dY(p) * X_bh(p + t - r)        (summed over the batch)
 * in gather form: a flipped-mask stencil over dY whose border pixels add the
 * halo gradients their boundary mode folds onto them (CLAMP, MIRROR, REPEAT;
 * CONSTANT halos take no gradient but const_val enters dw; UNDEFINED reads
 * are clamped to the image, so it folds as CLAMP).  dw is a fixed-order
 * reduction (deterministic).  F32 images; in, grad_out and grad_in have the
 * same size and batch.  No counterpart in the reference (it has no adjoints).
 * The flipped-mask stencil over the image runs through the forward's
 * kernels (hipacc_hip_local, CONSTANT 0); the folded border terms are
 * evaluated once per halo position and added to the border pixels.
 * workspace: NULL (the library's per-device buffer, grown on demand) or
 * device memory of hipacc_hip_local_grad_workspace_bytes(desc) bytes.       */
typedef struct hipacc_local_grad_desc {
  hipacc_view in;            /* X, the forward input                        */
  hipacc_view grad_out;      /* dL/dY                                       */
  hipacc_view grad_in;       /* dL/dX, overwritten; data NULL = skip        */
  int32_t mask_w, mask_h;    /* odd, 1..HIPACC_MAX_MASK_DIM                 */
  const float* weights_f32;  /* host mask, or                               */
  const float* weights_dev;  /* device mask (takes precedence)              */
  float*  grad_weights_dev;  /* device, mask_h*mask_w floats, overwritten;
                                NULL = skip                                 */
  int32_t boundary;
  double  const_val;
  void*   workspace;
} hipacc_local_grad_desc;
size_t hipacc_hip_local_grad_workspace_bytes(const hipacc_local_grad_desc* desc);
int hipacc_hip_local_grad(const hipacc_local_grad_desc* desc, void* stream);
/* The dw kernel variant: 0 automatic, 1 LDS tiles, 2 row streaming (square
 * masks up to 9x9; other shapes use a generic kernel).  All variants give
 * the same dX; dw agrees to rounding (each sums in its own fixed order).    */
int hipacc_hip_local_grad_variant(const hipacc_local_grad_desc* desc, int32_t variant,
                                  void* stream);

/* Allocation-free form of hipacc_hip_local_grad: the same kernels and
 * results, with the workspace taken from a static device array of
 * HIPACC_NOALLOC_SCRATCH_BYTES that the library holds on each device until
 * the process ends.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory -- code under HIP graph capture, tight
 * training or serving loops, harnesses that forbid hipMalloc.  It calls no
 * allocator.  desc->workspace must be NULL.  A descriptor whose
 * hipacc_hip_local_grad_workspace_bytes exceeds the array fails with
 * HIPACC_ERR_INVALID_ARG before anything is launched.  Every call on a
 * device shares that device's array: calls that overlap on one device
 * (several streams or host threads) race.  Launches are not timed by
 * hipacc_hip_set_timing (no events are created).                           */
#define HIPACC_NOALLOC_SCRATCH_BYTES (16 * 1024 * 1024)
int hipacc_hip_local_grad_noalloc(const hipacc_local_grad_desc* desc, void* stream);

/* ---- bilateral filter with a learnable range sigma -------------------- */
/* Training forward: hipacc_hip_bilateral (F32, whole images) that also
 * writes the normaliser D(p) = SUM_t s_t, s_t = exp(-c*(v_t - X(p))^2) *
 * spatial[t], c = 0.5/sigma_r^2, so Y(p) = SUM_t s_t*v_t / D(p).
 * Backward: given dY, with a(p) = dY(p)/D(p),
 *   dX(q) = 2c*a(q) * SUM_t s_{q,t} (v_t - X(q)) (v_t - Y(q))         (centre)
 *         + SUM_{p,t: bh(p+t-r) = q} a(p) s_{p,t} [1 - 2c (X(q) - X(p)) (X(q) - Y(p))]
 *   dL/dsigma_r = (-1/sigma_r^3) * dL/dc,
 *   dL/dc = -SUM_p a(p) SUM_t s_{p,t} (v_t - X(p))^2 (v_t - Y(p)),
 * recomputing the range weights (nothing per tap is stored); the second dX
 * sum is gathered per q with the boundary folding of hipacc_hip_local_grad;
 * the range factor exp(-c*d^2) of a pixel pair is computed once and serves
 * both the centre term of q and the neighbour term of p.  No counterpart in
 * the reference.                                                            */
typedef struct hipacc_bilateral_grad_desc {
  hipacc_view in;            /* X                                           */
  hipacc_view out;           /* Y (written by train, read by grad)          */
  hipacc_view norm;          /* D (written by train, read by grad)          */
  hipacc_view grad_out;      /* dL/dY (grad)                                */
  hipacc_view grad_in;       /* dL/dX (grad, overwritten); data NULL = skip */
  int32_t ksize;             /* odd, 1..HIPACC_MAX_MASK_DIM                 */
  float   sigma_d;
  float   sigma_r;
  const float* spatial;      /* host table or NULL (from sigma_d)           */
  int32_t boundary;
  double  const_val;
  float*  grad_sigma_r_dev;  /* device, 1 float, overwritten; NULL = skip   */
  void*   workspace;         /* NULL or hipacc_hip_bilateral_grad_workspace_bytes() */
} hipacc_bilateral_grad_desc;
size_t hipacc_hip_bilateral_grad_workspace_bytes(const hipacc_bilateral_grad_desc* desc);
int hipacc_hip_bilateral_train(const hipacc_bilateral_grad_desc* desc, void* stream);
int hipacc_hip_bilateral_grad(const hipacc_bilateral_grad_desc* desc, void* stream);

/* ---- resampling point operators (Interpolate::NN / LF) ---------------- */
/* Per pixel (x, y) of the iteration space `out`, the accessor `in` (an F32
 * window of w_in x h_in) is read at the mapped position of Interpolation::
 * interpolate (dsl/image.hpp:390-422), in float:
 *     s = w_in / (float)out.width;  xm = s/2 + s*x;   (same for y)
 *   NN: v = in_bh((int)xm, (int)ym)
 *   LF: xb = max(xm - 0.5f, 0), xi = (int)xb, fx = xb - xi (same for y)
 *       v = (1-fx)(1-fy) in_bh(xi,yi) + fx(1-fy) in_bh(xi+1,yi)
 *         + (1-fx) fy in_bh(xi,yi+1) + fx fy in_bh(xi+1,yi+1)
 *       (products and sums left to right, no contraction)
 * and combined with `in2` (F32, read at (x, y) of its window, no boundary):
 *   COPY out = v     SUB out = in2 - v     ADD out = v + in2
 * These are the Subsample, DifferenceOfGaussian and Restore kernels of the
 * pyramid sample (samples-public/5_Other/Gaussian_Laplacian_Pyramid/src/
 * main.cpp:67-117).  in_bh uses `boundary` over the accessor window.        */
typedef enum { HIPACC_INTERP_NN = 1, HIPACC_INTERP_LF = 2, HIPACC_INTERP_B5 = 3,
               HIPACC_INTERP_CF = 4, HIPACC_INTERP_L3 = 5 } hipacc_interp;
typedef enum { HIPACC_RESAMPLE_COPY = 0, HIPACC_RESAMPLE_SUB = 1,
               HIPACC_RESAMPLE_ADD = 2 } hipacc_resample_op;
typedef struct hipacc_resample_desc {
  hipacc_view in;            /* accessor read through the interpolation     */
  hipacc_view in2;           /* second operand (SUB / ADD), out's size      */
  hipacc_view out;           /* iteration space                             */
  int32_t interp;            /* hipacc_interp                               */
  int32_t op;                /* hipacc_resample_op                          */
  int32_t boundary;
  double  const_val;
  int32_t xf, yf;            /* the accessor call in(xf, yf): pixel (x, y) of
                                the iteration space maps x + xf, y + yf     */
} hipacc_resample_desc;
int hipacc_hip_resample(const hipacc_resample_desc* desc, void* stream);

/* ---- batched Gaussian / Laplacian pyramids ---------------------------- */
/* Level sizes halve with truncation (dsl/pyramid.hpp:103-114).  For level
 * l -> l+1 of every image (the pyramid sample's traversal body,
 * samples-public/5_Other/Gaussian_Laplacian_Pyramid/src/main.cpp:196-223):
 *   G[l+1] = NN-subsample( conv5x5(G[l], mask, boundary) )
 *   L[l]   = G[l] - LF-upsample(G[l+1])          (LF accessor: CLAMP)
 * with the convolution as hipacc_hip_local (float SUM, row-major taps,
 * fp_contract as there).  L[levels-1] is G[levels-1] itself (not written).
 * A plan holds, for all images, a work list of tiles per level (128x32 for
 * the large levels, 64 x tile_rows for the small ones) and executes each
 * level as one launch over that list, one block per tile (the blur is
 * evaluated only at the subsampled points, the upsampled G[l+1] comes from
 * LDS); the plan's launches are captured once into a hipGraph and replayed
 * by hipacc_hip_pyramid_run.  Results equal the three-kernels-per-level
 * sequence of the sample, run per image with hipacc_hip_local and
 * hipacc_hip_resample, bit for bit.
 * Replaces per-level kernel launches driven by HipaccPyramidTraversor
 * (runtime/hipacc_base_standalone.hpp:85-254).                              */
#define HIPACC_MAX_LEVELS 16
typedef struct hipacc_pyramid_image {
  int32_t width, height;                 /* level 0                          */
  float*  gauss[HIPACC_MAX_LEVELS];      /* G[l], device; G[0] is the input  */
  int32_t gauss_stride[HIPACC_MAX_LEVELS];
  float*  lap[HIPACC_MAX_LEVELS];        /* L[l], device, l < levels-1       */
  int32_t lap_stride[HIPACC_MAX_LEVELS];
} hipacc_pyramid_image;
typedef struct hipacc_pyramid_desc {
  int32_t num_images;
  int32_t levels;                        /* 2..HIPACC_MAX_LEVELS             */
  const hipacc_pyramid_image* images;    /* host array of num_images         */
  const float* mask;                     /* host 5x5, NULL: (1 4 6 4 1)^T (1 4 6 4 1) / 256 */
  int32_t boundary;                      /* of the convolution (CONSTANT: 0) */
  int32_t fp_contract;
  int32_t tile_rows;                     /* rows of L per narrow tile (small levels):
                                            0 = default (16), 8, 16, 32     */
} hipacc_pyramid_desc;
typedef struct hipacc_pyramid_plan hipacc_pyramid_plan;
/* Builds the work lists (device memory, uploaded on `stream`, which is
 * synchronised) and captures the level launches into a graph.  The image
 * buffers must stay allocated while the plan lives; their contents may
 * change between runs.                                                      */
int hipacc_hip_pyramid_plan_create(const hipacc_pyramid_desc* desc, void* stream,
                                   hipacc_pyramid_plan** plan);
/* Replays the captured graph on `stream` (asynchronous).                   */
int hipacc_hip_pyramid_run(hipacc_pyramid_plan* plan, void* stream);
/* The same launches issued directly, without the graph.                    */
int hipacc_hip_pyramid_run_direct(hipacc_pyramid_plan* plan, void* stream);
int hipacc_hip_pyramid_plan_destroy(hipacc_pyramid_plan* plan);
/* Host only (no device call): builds the work lists that
 * hipacc_hip_pyramid_plan_create builds for these level-0 sizes, proves that
 * every tile's staged regions fit the LDS capacity of the kernel that runs
 * it (HIPACC_ERR_INVALID_ARG otherwise, exactly as plan_create fails), and
 * writes the narrow and wide tile count of each level (levels-1 entries;
 * either pointer may be NULL).  An inspection helper of the plan; the
 * reference's pyramids (runtime/hipacc_base.hpp:131-191) have no work list. */
int hipacc_hip_pyramid_plan_tiles(const int32_t* widths, const int32_t* heights,
                                  int32_t num_images, int32_t levels, int32_t tile_rows,
                                  int32_t* narrow_tiles, int32_t* wide_tiles);

/* ---- kernel variants and the autotuner ------------------------------- */
/* A local or separable operator runs one of these kernel variants:
 *   1 tile    LDS-staged tile + halo; param = output rows per wave (2/4/8)
 *   2 stream  row-streaming strips with a register prefetch ring and
 *             cross-lane halos; param = output rows per wave segment
 *   3 generic one pixel per thread, any shape and reduction mode
 *   4 wide    stream with 8 pixels per lane instead of 4
 *   5 ring    block row ring in LDS: local weighted SUM f32 -> f32 (square
 *             masks 3..9), u8 -> u8 and u8 -> i16/u32/i32 integer sums
 *             (3x3, 5x5); separable f32, u8, and integer sums of u8 pixels;
 *             param = output rows per block segment, 0 = sized so that
 *             every block is resident at once
 * All variants compute identical results.  The plain entry points use the
 * tuned variant of the operator's shape when the tune cache has one, else
 * ring (for 8-bit local operators only from 2977 columns up), else stream,
 * else tile, else generic.  The tuner runs every candidate once on the
 * caller's buffers (the output is overwritten), then times it back to back
 * over buffer sets of the operator's extents that together exceed the
 * Infinity Cache (512 MiB, allocated for the call; the caller's buffers
 * alone when one set is below 8 MiB), and records the fastest per (kind,
 * mask size, pixel types, accumulator, boundary, contraction, device
 * weights, size class = floor(log2(width * height * batch)) of the
 * iteration space).  The cache lives in the process and can be saved /
 * loaded.
 * Replaces the reference's fixed per-device kernel configuration
 * (lib/DSL/ClassRepresentation.cpp:256-489 setDefaultConfig, -use-config /
 * -pixels-per-thread, include/hipacc/Device/TargetDescription.h:64-180).   */
typedef struct hipacc_tune_entry {
  int32_t kind;        /* 0 local weighted SUM, 1 separable, 2 box sum         */
  int32_t kx, ky;
  int32_t in_dtype, out_dtype, acc_dtype, boundary, fp_contract, weights_dev;
  int32_t size_class;  /* floor(log2(width * height * batch))                 */
  int32_t variant;     /* 1 tile, 2 stream, 3 generic, 4 wide stream, 5 ring  */
  int32_t param;
  float   ms;          /* measured time per launch                            */
} hipacc_tune_entry;
/* variant 0 = automatic choice; a forced variant without a kernel for the
 * operator's shape returns HIPACC_ERR_UNSUPPORTED.                           */
int hipacc_hip_local_variant(const hipacc_local_desc* desc, int32_t variant, int32_t param,
                             void* stream);
int hipacc_hip_separable_variant(const hipacc_separable_desc* desc, int32_t variant,
                                 int32_t param, void* stream);
int hipacc_hip_tune_local(const hipacc_local_desc* desc, int32_t iters, hipacc_tune_entry* best,
                          void* stream);
int hipacc_hip_tune_separable(const hipacc_separable_desc* desc, int32_t iters,
                              hipacc_tune_entry* best, void* stream);
int hipacc_hip_tune_set(const hipacc_tune_entry* entry);
int hipacc_hip_tune_save(const char* path);
int hipacc_hip_tune_load(const char* path);

/* ---- library / memory ------------------------------------------------- */
/* hipaccInitCUDA (runtime/hipacc_cu_standalone.hpp:112-160) analogue:
 * selects `device`, checks it is gfx950, creates the per-device state.     */
int hipacc_hip_init(int device);
const char* hipacc_hip_strerror(int status);
const char* hipacc_hip_last_error(void);
/* hipaccCreateMemory (runtime/hipacc_cu.tpp:33-83): pitched allocation, row
 * pitch rounded up to `alignment` bytes (0 = 256); pixels of 1, 2, 4, 8 or
 * 16 bytes (the 4-wide pixel types up to float4 / int4).                    */
int hipacc_hip_alloc2d(void** ptr, int32_t* stride_elems, int32_t width,
                       int32_t height, int32_t elem_bytes, int32_t alignment);
int hipacc_hip_free(void* ptr);
/* hipaccWriteMemory / hipaccReadMemory (runtime/hipacc_cu.tpp:85-193):
 * host rows are dense (width elements); blocking.                           */
int hipacc_hip_upload2d(void* dst_dev, int32_t dst_stride_elems,
                        const void* src_host, int32_t width, int32_t height,
                        int32_t elem_bytes, void* stream);
int hipacc_hip_download2d(void* dst_host, const void* src_dev,
                          int32_t src_stride_elems, int32_t width,
                          int32_t height, int32_t elem_bytes, void* stream);
/* hipaccCopyMemoryRegion (runtime/hipacc_cu_standalone.hpp:166-216):
 * device-to-device copy of a window, asynchronous.                         */
int hipacc_hip_copy_region(const hipacc_view* src, const hipacc_view* dst,
                           void* stream);
/* Fills `count` elements of elem_bytes (1, 2 or 4) at dst_dev with the low
 * bytes of `value`, with a kernel (safe inside stream capture).            */
int hipacc_hip_fill(void* dst_dev, uint32_t value, size_t count,
                    int32_t elem_bytes, void* stream);
/* SplitMix64 synthetic pixels, counter form: element i of the image holds
 * the (i+1)-th output of SplitMix64 seeded with `seed` (dense row-major
 * numbering over width*height, written at the view's pitch).  U8: x >> 56;
 * F32: (x >> 40) * 2^-24.  Used by tests and the benchmark.                */
int hipacc_hip_fill_random(const hipacc_view* img, uint64_t seed, void* stream);

/* hipacc_last_kernel_timing (runtime/hipacc_base.hpp:54-66): when timing is
 * enabled, each entry point brackets its launches with hipEvents on the
 * stream and synchronises, and the elapsed milliseconds are kept here.  The
 * switch and the last time are per host thread, as the reference's timing
 * singleton is thread_local (runtime/hipacc_base_standalone.hpp:35-39).  A
 * call that fails leaves the last time unchanged.                          */
int   hipacc_hip_set_timing(int enabled);
float hipacc_hip_last_kernel_timing(void);
/* The calling thread's timing switch, and a time recorded by a launch made
 * outside the library (kernels the C++ runtime compiles from a program's
 * own kernel() code), so that hipacc_last_kernel_timing() reports it:
 * HipaccKernelTimingBase::set_timing (runtime/hipacc_base.hpp:54-66).      */
int   hipacc_hip_get_timing(void);
int   hipacc_hip_set_last_kernel_timing(float ms);

/* Library version string, e.g. "hipacc-gfx950 0.1". */
const char* hipacc_hip_version(void);

#ifdef __cplusplus
}
#endif
#endif /* HIPACC_HIP_H */
