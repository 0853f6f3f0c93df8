// capi.hip -- the extern "C" entry points of libhipacc_gfx950.so: argument
// validation, descriptor -> kernel-argument translation, per-device state,
// timing and error reporting.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "hh_fused.hpp"
#include "hh_grad.hpp"
#include "hh_pyramid.hpp"
#include "hh_tile.hpp"

namespace hh {

// ---------------------------------------------------------------------------
// Errors
// ---------------------------------------------------------------------------
static thread_local std::string t_last_error;

int set_error(int status, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  t_last_error = buf;
  return status;
}

int hip_status(hipError_t e, const char* what) {
  if (e == hipSuccess) return HIPACC_OK;
  return set_error(e == hipErrorOutOfMemory ? HIPACC_ERR_OUT_OF_MEMORY : HIPACC_ERR_HIP,
                   "%s: %s (%d)", what, hipGetErrorString(e), (int)e);
}

// ---------------------------------------------------------------------------
// Per-device state: the reduction workspace.
// ---------------------------------------------------------------------------
constexpr int kMaxDevices = 64;
constexpr size_t kRedWorkspace = 1024 * sizeof(unsigned long long) + 256;

struct DeviceState {
  std::mutex mu;
  bool ready = false;
  void* red_ws = nullptr;
  void* grad_ws = nullptr;   // adjoint workspace, allocated / grown on demand
  size_t grad_bytes = 0;
  void* scratch_ptr = nullptr;   // this device's copy of `scratch`
};
static DeviceState g_dev[kMaxDevices];

// Workspace of the allocation-free adjoint (hipacc_hip_local_grad_noalloc):
// a static device array, one copy per device, alive until the process ends.
// Host code never passes `scratch` itself to a kernel (on the host that name
// is only a stand-in); it uses the device address that hipGetSymbolAddress
// returns for the current device, kept in DeviceState::scratch_ptr.
static __device__ __align__(256) unsigned char scratch[HIPACC_NOALLOC_SCRATCH_BYTES];

// The current device's address of `scratch`, looked up once per device
// (with that device current).
static int scratch_address(int dev, void** p) {
  DeviceState& d = g_dev[dev];
  std::lock_guard<std::mutex> lk(d.mu);
  if (!d.scratch_ptr) {
    void* q = nullptr;
    const hipError_t e = hipGetSymbolAddress(&q, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_status(e, "hipGetSymbolAddress(scratch)");
    if (!q) return set_error(HIPACC_ERR_HIP, "hipGetSymbolAddress(scratch) returned NULL");
    d.scratch_ptr = q;
  }
  *p = d.scratch_ptr;
  return HIPACC_OK;
}

static int current_device(int* dev) {
  hipError_t e = hipGetDevice(dev);
  if (e != hipSuccess) return hip_status(e, "hipGetDevice");
  if (*dev < 0 || *dev >= kMaxDevices) return set_error(HIPACC_ERR_NO_DEVICE, "device %d", *dev);
  return HIPACC_OK;
}

static int ensure_workspace(int dev, void** ws) {
  DeviceState& d = g_dev[dev];
  std::lock_guard<std::mutex> lk(d.mu);
  if (!d.red_ws) {
    hipError_t e = hipMalloc(&d.red_ws, kRedWorkspace);
    if (e != hipSuccess) return hip_status(e, "hipMalloc(reduction workspace)");
    e = hipMemset(d.red_ws, 0, kRedWorkspace);
    if (e != hipSuccess) return hip_status(e, "hipMemset(reduction workspace)");
    e = hipDeviceSynchronize();
    if (e != hipSuccess) return hip_status(e, "hipDeviceSynchronize");
  }
  *ws = d.red_ws;
  return HIPACC_OK;
}

// Per-stream workspaces (the histogram's partial bins): the calls on one
// stream run in order, so a stream's buffer is never used by two calls at
// once, and calls on different streams (or host threads with their own
// streams) get different buffers.  Fixed size, never freed before the
// process ends.
constexpr size_t kStreamWorkspace = sizeof(uint32_t) * kHistPartialEntries;
static std::mutex g_stream_ws_mu;
static std::map<std::pair<int, hipStream_t>, void*> g_stream_ws;

static int stream_workspace(hipStream_t s, void** ws) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_status(e, "hipGetDevice");
  std::lock_guard<std::mutex> lk(g_stream_ws_mu);
  void*& p = g_stream_ws[std::make_pair(dev, s)];
  if (!p) {
    e = hipMalloc(&p, kStreamWorkspace);
    if (e != hipSuccess) {
      p = nullptr;
      return hip_status(e, "hipMalloc(stream workspace)");
    }
  }
  *ws = p;
  return HIPACC_OK;
}

// The adjoint workspace of the current device, at least `bytes` long.
// Growing it frees the old buffer after the device is idle.
static int ensure_grad_workspace(size_t bytes, void** ws) {
  int dev = 0;
  int st = current_device(&dev);
  if (st) return st;
  DeviceState& d = g_dev[dev];
  std::lock_guard<std::mutex> lk(d.mu);
  if (d.grad_bytes < bytes) {
    if (d.grad_ws) {
      hipError_t e = hipDeviceSynchronize();
      if (e != hipSuccess) return hip_status(e, "hipDeviceSynchronize");
      (void)hipFree(d.grad_ws);
      d.grad_ws = nullptr;
      d.grad_bytes = 0;
    }
    hipError_t e = hipMalloc(&d.grad_ws, bytes);
    if (e != hipSuccess) return hip_status(e, "hipMalloc(gradient workspace)");
    d.grad_bytes = bytes;
  }
  *ws = d.grad_ws;
  return HIPACC_OK;
}

// ---------------------------------------------------------------------------
// Timing (hipacc_last_kernel_timing)
// ---------------------------------------------------------------------------
// Per host thread, as the reference's timing singleton
// (runtime/hipacc_base_standalone.hpp:35-39): a thread that enables timing
// times its own launches and reads its own last time.
static thread_local bool g_timing = false;
static thread_local float g_last_ms = 0.0f;

struct Timer {
  hipStream_t s;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  explicit Timer(hipStream_t st, bool enabled = true) : s(st) {
    if (enabled && g_timing && hipEventCreate(&e0) == hipSuccess &&
        hipEventCreate(&e1) == hipSuccess)
      hipEventRecord(e0, s);
  }
  int finish(int st) {
    if (e0 && e1) {
      if (st == HIPACC_OK) {
        hipEventRecord(e1, s);
        hipEventSynchronize(e1);
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, e0, e1) == hipSuccess) g_last_ms = ms;
      }
      hipEventDestroy(e0);
      hipEventDestroy(e1);
      e0 = e1 = nullptr;
    }
    return st;
  }
  ~Timer() {
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
  }
};

// ---------------------------------------------------------------------------
// Descriptor checks
// ---------------------------------------------------------------------------
static bool valid_dtype(int dt) { return dt >= HIPACC_U8 && dt <= HIPACC_F32X4; }

// The scalar view of a vector-pixel view: its channels as an image of 4x
// the columns (elementwise operators only).
static hipacc_view channel_view(const hipacc_view& v) {
  hipacc_view c = v;
  c.dtype = dtype_base(v.dtype);
  c.stride *= 4;
  c.img_width *= 4;
  c.width *= 4;
  c.offset_x *= 4;
  c.plane_stride *= 4;
  return c;
}

static int check_view(const hipacc_view& v, const char* name) {
  if (!v.data) return set_error(HIPACC_ERR_INVALID_ARG, "%s: data is NULL", name);
  if (!valid_dtype(v.dtype)) return set_error(HIPACC_ERR_INVALID_ARG, "%s: dtype %d", name, v.dtype);
  if (v.img_width <= 0 || v.img_height <= 0 || v.stride < v.img_width)
    return set_error(HIPACC_ERR_INVALID_ARG, "%s: image %dx%d stride %d", name, v.img_width,
                     v.img_height, v.stride);
  if (v.width <= 0 || v.height <= 0 || v.offset_x < 0 || v.offset_y < 0 ||
      v.offset_x + v.width > v.img_width || v.offset_y + v.height > v.img_height)
    return set_error(HIPACC_ERR_INVALID_ARG, "%s: window %dx%d+%d+%d outside image %dx%d", name,
                     v.width, v.height, v.offset_x, v.offset_y, v.img_width, v.img_height);
  if (v.batch < 0) return set_error(HIPACC_ERR_INVALID_ARG, "%s: batch %d", name, v.batch);
  if (v.batch > 1 && v.plane_stride < (int64_t)v.stride * v.img_height)
    return set_error(HIPACC_ERR_INVALID_ARG, "%s: plane_stride %lld too small", name,
                     (long long)v.plane_stride);
  return HIPACC_OK;
}

static int batch_of(const hipacc_view& v) { return v.batch > 1 ? v.batch : 1; }

// Whether every row start of the window is aligned to `align` bytes.
static bool rows_aligned(const hipacc_view& v, int align) {
  const int eb = dtype_bytes(v.dtype);
  const uintptr_t p = reinterpret_cast<uintptr_t>(v.data);
  if (p % align) return false;
  if (((long long)v.stride * eb) % align) return false;
  if (((long long)v.offset_x * eb) % align) return false;
  if (batch_of(v) > 1 && (v.plane_stride * eb) % align) return false;
  return true;
}

// An F32 image viewed whole (the adjoint kernels' operands).
static int grad_image(const hipacc_view& v, const char* name) {
  int st = check_view(v, name);
  if (st) return st;
  if (v.dtype != HIPACC_F32) return set_error(HIPACC_ERR_UNSUPPORTED, "%s: F32 images only", name);
  if (v.offset_x || v.offset_y || v.width != v.img_width || v.height != v.img_height)
    return set_error(HIPACC_ERR_UNSUPPORTED, "%s: the window must be the whole image", name);
  return HIPACC_OK;
}
static bool same_shape(const hipacc_view& a, const hipacc_view& b) {
  return a.width == b.width && a.height == b.height && batch_of(a) == batch_of(b);
}

static void const_views(int dtype, double cv, float* cf, int* ci) {
  switch (dtype) {
    case HIPACC_U8: *ci = (uint8_t)(long long)cv; break;
    case HIPACC_I8: *ci = (int8_t)(long long)cv; break;
    case HIPACC_U16: *ci = (uint16_t)(long long)cv; break;
    case HIPACC_I16: *ci = (int16_t)(long long)cv; break;
    case HIPACC_U32: *ci = (int32_t)(uint32_t)(long long)cv; break;
    case HIPACC_I32: *ci = (int32_t)(long long)cv; break;
    default:
      *cf = (float)cv;
      *ci = (int32_t)*cf;
      return;
  }
  *cf = dtype == HIPACC_U32 ? (float)(uint32_t)*ci : (float)*ci;
}

static int make_geom(Geom& g, const hipacc_view& in, const hipacc_view& out, int boundary,
                     double const_val, int cvt, float post_div) {
  int st = check_view(in, "in");
  if (st) return st;
  if ((st = check_view(out, "out"))) return st;
  if (batch_of(in) != batch_of(out))
    return set_error(HIPACC_ERR_INVALID_ARG, "batch of in (%d) and out (%d) differ", batch_of(in),
                     batch_of(out));
  if (boundary < HIPACC_BOUNDARY_UNDEFINED || boundary > HIPACC_BOUNDARY_CONSTANT)
    return set_error(HIPACC_ERR_INVALID_ARG, "boundary %d", boundary);
  if (cvt != HIPACC_CVT_TRUNC && cvt != HIPACC_CVT_ROUND)
    return set_error(HIPACC_ERR_INVALID_ARG, "cvt %d", cvt);
  std::memset(&g, 0, sizeof(g));
  g.in = in.data;
  g.out = out.data;
  g.in_plane = in.plane_stride;
  g.out_plane = out.plane_stride;
  g.in_stride = in.stride;
  g.out_stride = out.stride;
  g.in_dtype = in.dtype;
  g.out_dtype = out.dtype;
  g.in_imgw = in.img_width;
  g.in_imgh = in.img_height;
  g.lo_x = in.offset_x;
  g.lo_y = in.offset_y;
  g.hi_x = in.offset_x + in.width;
  g.hi_y = in.offset_y + in.height;
  g.in_ox = in.offset_x;
  g.in_oy = in.offset_y;
  g.out_ox = out.offset_x;
  g.out_oy = out.offset_y;
  g.is_w = out.width;
  g.is_h = out.height;
  g.batch = batch_of(in);
  g.boundary = boundary;
  const_views(in.dtype, const_val, &g.const_f, &g.const_i);
  g.cvt = cvt;
  g.post_div = post_div;
  return HIPACC_OK;
}

}  // namespace hh

using namespace hh;

extern "C" {

const char* hipacc_hip_version(void) { return "hipacc-gfx950 0.1"; }

const char* hipacc_hip_strerror(int status) {
  switch (status) {
    case HIPACC_OK: return "ok";
    case HIPACC_ERR_INVALID_ARG: return "invalid argument";
    case HIPACC_ERR_UNSUPPORTED: return "unsupported";
    case HIPACC_ERR_HIP: return "HIP runtime error";
    case HIPACC_ERR_NO_DEVICE: return "no usable device";
    case HIPACC_ERR_OUT_OF_MEMORY: return "out of device memory";
  }
  return "unknown status";
}

const char* hipacc_hip_last_error(void) { return t_last_error.c_str(); }

int hipacc_hip_init(int device) {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0) return set_error(HIPACC_ERR_NO_DEVICE, "no HIP device");
  if (device < 0 || device >= n || device >= kMaxDevices)
    return set_error(HIPACC_ERR_NO_DEVICE, "device %d of %d", device, n);
  if ((e = hipSetDevice(device)) != hipSuccess) return hip_status(e, "hipSetDevice");
  hipDeviceProp_t prop;
  if ((e = hipGetDeviceProperties(&prop, device)) != hipSuccess)
    return hip_status(e, "hipGetDeviceProperties");
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return set_error(HIPACC_ERR_NO_DEVICE, "device %d is %s, this library is built for gfx950",
                     device, prop.gcnArchName);
  void* ws = nullptr;
  int st = ensure_workspace(device, &ws);
  if (st) return st;
  if ((st = scratch_address(device, &ws))) return st;
  g_dev[device].ready = true;
  return HIPACC_OK;
}

int hipacc_hip_set_timing(int enabled) {
  g_timing = enabled != 0;
  return HIPACC_OK;
}
float hipacc_hip_last_kernel_timing(void) { return g_last_ms; }
int hipacc_hip_get_timing(void) { return g_timing ? 1 : 0; }
int hipacc_hip_set_last_kernel_timing(float ms) {
  if (!(ms >= 0.0f)) return set_error(HIPACC_ERR_INVALID_ARG, "kernel timing %f ms", (double)ms);
  g_last_ms = ms;
  return HIPACC_OK;
}

}  // extern "C"

namespace hh {

// Translates a local descriptor.  *box is set when the operator is a box sum
// that the separable running-sum kernel computes exactly (then sa is set).
static int prep_local(const hipacc_local_desc* d, LocalArgs& a, int& accf, bool& box, SepArgs& sa) {
  if (!d) return set_error(HIPACC_ERR_INVALID_ARG, "local: desc is NULL");
  std::memset(&a, 0, sizeof(a));
  int st = make_geom(a.g, d->in, d->out, d->boundary, d->const_val, d->cvt, d->post_div);
  if (st) return st;
  if (d->mask_w < 1 || d->mask_h < 1 || d->mask_w > HIPACC_MAX_MASK_DIM ||
      d->mask_h > HIPACC_MAX_MASK_DIM || !(d->mask_w & 1) || !(d->mask_h & 1))
    return set_error(HIPACC_ERR_INVALID_ARG, "local: mask %dx%d (odd, <= %d)", d->mask_w,
                     d->mask_h, HIPACC_MAX_MASK_DIM);
  if (d->acc_dtype != HIPACC_F32 && d->acc_dtype != HIPACC_I32)
    return set_error(HIPACC_ERR_INVALID_ARG, "local: acc_dtype must be F32 or I32");
  if (d->reduce < HIPACC_REDUCE_SUM || d->reduce > HIPACC_REDUCE_PROD)
    return set_error(HIPACC_ERR_INVALID_ARG, "local: reduce %d", d->reduce);
  const bool vecpx = dtype_is_vec(d->in.dtype);
  if (vecpx != dtype_is_vec(d->out.dtype))
    return set_error(HIPACC_ERR_INVALID_ARG, "local: vector and scalar pixels mixed");
  const int n = d->mask_w * d->mask_h;
  accf = d->acc_dtype == HIPACC_F32;
  if (d->use_weights) {
    if (d->weights_dev) {
      if (!accf) return set_error(HIPACC_ERR_INVALID_ARG, "local: weights_dev needs acc_dtype F32");
    } else if (accf) {
      if (d->weights_f32) {
        std::memcpy(a.w.f, d->weights_f32, n * sizeof(float));
      } else if (d->weights_i32) {
        for (int t = 0; t < n; ++t) a.w.f[t] = (float)d->weights_i32[t];
      } else {
        return set_error(HIPACC_ERR_INVALID_ARG, "local: use_weights without weights");
      }
    } else {
      if (d->weights_i32) {
        std::memcpy(a.w.i, d->weights_i32, n * sizeof(int32_t));
      } else if (d->weights_f32) {
        for (int t = 0; t < n; ++t) a.w.i[t] = (int32_t)d->weights_f32[t];
      } else {
        return set_error(HIPACC_ERR_INVALID_ARG, "local: use_weights without weights");
      }
    }
  }
  bool any = false;
  for (int t = 0; t < n; ++t) {
    a.domain[t] = d->domain ? (d->domain[t] != 0) : 1;
    any = any || a.domain[t];
  }
  if (!any) return set_error(HIPACC_ERR_INVALID_ARG, "local: empty domain");
  a.mask_w = d->mask_w;
  a.mask_h = d->mask_h;
  a.reduce = d->reduce;
  a.use_weights = d->use_weights != 0;
  a.fma = d->fp_contract != 0;
  a.wdev = d->use_weights ? d->weights_dev : nullptr;
  const int ebi = dtype_bytes(d->in.dtype), ebo = dtype_bytes(d->out.dtype);
  const int vec = 16 / ebi;
  a.in_vec_ok = rows_aligned(d->in, 16) && (d->in.offset_x % vec) == 0;
  a.out_vec_ok = rows_aligned(d->out, std::min(16, vec * ebo));
  // An integer SUM over a full all-ones window (a box sum) is the outer
  // product of two all-ones rows; integer sums are exact in any order, so
  // the separable running-sum kernels compute the same values.  CONSTANT
  // qualifies only with the constant 0 (otherwise the border rows differ).
  box = !vecpx && !accf && a.reduce == HIPACC_REDUCE_SUM && !a.wdev && d->in.width == d->out.width &&
        d->in.height == d->out.height && a.mask_w <= 2 * HIPACC_MAX_SEP_RADIUS + 1 &&
        (a.g.boundary != HIPACC_BOUNDARY_CONSTANT || a.g.const_i == 0);
  for (int t = 0; box && t < n; ++t) box = a.domain[t] && (!a.use_weights || a.w.i[t] == 1);
  if (box) {
    std::memset(&sa, 0, sizeof(sa));
    sa.g = a.g;
    sa.kx = a.mask_w;
    sa.ky = a.mask_h;
    for (int i = 0; i < sa.kx; ++i) sa.rowi[i] = 1;
    for (int i = 0; i < sa.ky; ++i) sa.coli[i] = 1;
    sa.box = 1;
    sa.in_vec_ok = a.in_vec_ok;
    sa.out_vec_ok = a.out_vec_ok;
  }
  return HIPACC_OK;
}

static int prep_separable(const hipacc_separable_desc* d, SepArgs& a, int& accf) {
  if (!d) return set_error(HIPACC_ERR_INVALID_ARG, "separable: desc is NULL");
  std::memset(&a, 0, sizeof(a));
  int st = make_geom(a.g, d->in, d->out, d->boundary, d->const_val, d->cvt, d->post_div);
  if (st) return st;
  if (dtype_is_vec(d->in.dtype) || dtype_is_vec(d->out.dtype))
    return set_error(HIPACC_ERR_UNSUPPORTED,
                     "separable: vector pixels (use hipacc_hip_local with the 2D mask)");
  const int kmax = 2 * HIPACC_MAX_SEP_RADIUS + 1;
  if (d->ksize_x < 1 || d->ksize_y < 1 || d->ksize_x > kmax || d->ksize_y > kmax ||
      !(d->ksize_x & 1) || !(d->ksize_y & 1))
    return set_error(HIPACC_ERR_INVALID_ARG, "separable: ksize %dx%d", d->ksize_x, d->ksize_y);
  if (d->in.width != d->out.width || d->in.height != d->out.height)
    return set_error(HIPACC_ERR_UNSUPPORTED,
                     "separable: iteration space must have the accessor window's size");
  accf = d->acc_dtype == HIPACC_F32;
  if (!accf && d->acc_dtype != HIPACC_I32)
    return set_error(HIPACC_ERR_INVALID_ARG, "separable: acc_dtype must be F32 or I32");
  a.kx = d->ksize_x;
  a.ky = d->ksize_y;
  a.box = 0;
  if (accf) {
    if (!d->row_f32 || !d->col_f32) return set_error(HIPACC_ERR_INVALID_ARG, "separable: float weights missing");
    std::memcpy(a.rowf, d->row_f32, a.kx * sizeof(float));
    std::memcpy(a.colf, d->col_f32, a.ky * sizeof(float));
  } else {
    if (!d->row_i32 || !d->col_i32) return set_error(HIPACC_ERR_INVALID_ARG, "separable: integer weights missing");
    std::memcpy(a.rowi, d->row_i32, a.kx * sizeof(int32_t));
    std::memcpy(a.coli, d->col_i32, a.ky * sizeof(int32_t));
    bool ones = true;
    for (int i = 0; i < a.kx; ++i) ones = ones && a.rowi[i] == 1;
    for (int i = 0; i < a.ky; ++i) ones = ones && a.coli[i] == 1;
    a.box = ones;
  }
  const int ebi = dtype_bytes(d->in.dtype), ebo = dtype_bytes(d->out.dtype);
  const int vec = 16 / ebi;
  a.in_vec_ok = rows_aligned(d->in, 16) && (d->in.offset_x % vec) == 0;
  a.out_vec_ok = rows_aligned(d->out, std::min(16, vec * ebo));
  return HIPACC_OK;
}

static int run_local(const hipacc_local_desc* d, int variant, int param, hipStream_t s) {
  LocalArgs a;
  SepArgs sa;
  int accf = 0;
  bool box = false;
  int st = prep_local(d, a, accf, box, sa);
  if (st) return st;
  Timer tm(s);
  if (box) return tm.finish(launch_separable(sa, 0, 0, s, variant, param));
  return tm.finish(launch_local(a, accf, s, variant, param));
}

static int run_separable(const hipacc_separable_desc* d, int variant, int param, hipStream_t s) {
  SepArgs a;
  int accf = 0;
  int st = prep_separable(d, a, accf);
  if (st) return st;
  Timer tm(s);
  return tm.finish(launch_separable(a, accf, d->fp_contract != 0, s, variant, param));
}

// Buffer sets the tuner rotates over, so that each timed launch reads its
// input from HBM as it would in a pipeline over large images: copies of the
// caller's input / output extents (same pitch, same offset within 256 bytes,
// contents irrelevant to the timing) whose total reaches 512 MiB, twice the
// Infinity Cache.  Below 8 MiB per set the operator's working set stays in
// cache in real use as well, and the caller's buffers are used.
struct TuneRotation {
  char* base = nullptr;
  size_t in_bytes = 0, set = 0;
  uintptr_t in_off = 0, out_off = 0;
  int n = 0;
  ~TuneRotation() {
    if (base) (void)hipFree(base);
  }
  static size_t extent(const hipacc_view& v) {
    const long long nb = v.batch > 1 ? v.batch : 1;
    return (size_t)dtype_bytes(v.dtype) *
           ((size_t)v.plane_stride * (size_t)(nb - 1) + (size_t)v.stride * (size_t)v.img_height);
  }
  int init(const hipacc_view& in, const hipacc_view& out) {
    const size_t al = 256;
    in_bytes = (extent(in) + 2 * al - 1) / al * al;
    const size_t ob = (extent(out) + 2 * al - 1) / al * al;
    set = in_bytes + ob;
    if (set * 64 < ((size_t)512 << 20)) return HIPACC_OK;
    n = (int)((((size_t)512 << 20) + set - 1) / set);
    if (n < 2) n = 2;
    in_off = reinterpret_cast<uintptr_t>(in.data) % al;
    out_off = reinterpret_cast<uintptr_t>(out.data) % al;
    const hipError_t e = hipMalloc(&base, set * (size_t)n);
    if (e != hipSuccess) {
      base = nullptr;
      n = 0;
      return hip_status(e, "hipMalloc(tuner rotation buffers)");
    }
    return HIPACC_OK;
  }
  // Geometry of launch k: set k % n, or the caller's buffers without sets.
  void apply(Geom& g, int k) const {
    if (n == 0 || k < 0) return;
    char* p = base + (size_t)(k % n) * set;
    g.in = p + in_off;
    g.out = p + in_bytes + out_off;
  }
};

// Times every variant/parameter candidate of one operator and records the
// fastest in the tune cache.  run(variant, param, k) launches the candidate
// on rotation set k (k < 0: the caller's buffers).  Each candidate runs
// once on the caller's buffers (unsupported ones are skipped), once on every
// set, then `iters` timed launches back to back over the sets.
template <typename RunFn>
static int tune_candidates(RunFn run, const TuneKey& key, int iters, hipStream_t s,
                           hipacc_tune_entry* best, int nsets) {
  static const int cands[][2] = {
      {kVarStream, 8},      {kVarStream, 16},     {kVarStream, 24},     {kVarStream, 32},
      {kVarStream, 64},     {kVarStreamWide, 8},  {kVarStreamWide, 16}, {kVarStreamWide, 24},
      {kVarStreamWide, 32}, {kVarStreamWide, 64}, {kVarTile, 2},        {kVarTile, 4},
      {kVarTile, 8},        {kVarRing, 0},        {kVarRing, 32},       {kVarRing, 128},
      {kVarGeneric, 0}};
  if (iters < 1) iters = 1;
  hipEvent_t e0, e1;
  hipError_t e = hipEventCreate(&e0);
  if (e != hipSuccess) return hip_status(e, "hipEventCreate");
  if ((e = hipEventCreate(&e1)) != hipSuccess) {
    hipEventDestroy(e0);
    return hip_status(e, "hipEventCreate");
  }
  TuneVal bestv{0, 0, 0.0f};
  int st = HIPACC_OK;
  for (const auto& c : cands) {
    st = run(c[0], c[1], -1);
    if (st == HIPACC_ERR_UNSUPPORTED) {
      st = HIPACC_OK;
      continue;
    }
    if (st) break;
    for (int k = 0; k < nsets && st == HIPACC_OK; ++k) st = run(c[0], c[1], k);
    if (st) break;
    hipEventRecord(e0, s);
    for (int i = 0; i < iters && st == HIPACC_OK; ++i) st = run(c[0], c[1], nsets ? i % nsets : -1);
    hipEventRecord(e1, s);
    if (st) break;
    if ((e = hipEventSynchronize(e1)) != hipSuccess) {
      st = hip_status(e, "hipEventSynchronize");
      break;
    }
    float ms = 0.0f;
    hipEventElapsedTime(&ms, e0, e1);
    ms /= iters;
    if (bestv.variant == 0 || ms < bestv.ms) bestv = TuneVal{c[0], c[1], ms};
  }
  // the rotation buffers are freed by the caller after this point: the
  // stream must be idle
  if (st == HIPACC_OK && (e = hipStreamSynchronize(s)) != hipSuccess)
    st = hip_status(e, "hipStreamSynchronize");
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (st) return st;
  if (bestv.variant == 0) return set_error(HIPACC_ERR_UNSUPPORTED, "tune: no candidate ran");
  tune_store(key, bestv);
  if (best) {
    best->kind = key.kind;
    best->kx = key.kx;
    best->ky = key.ky;
    best->in_dtype = key.in_dt;
    best->out_dtype = key.out_dt;
    best->acc_dtype = key.acc_f ? HIPACC_F32 : HIPACC_I32;
    best->boundary = key.boundary;
    best->fp_contract = key.fma;
    best->weights_dev = key.wdev;
    best->size_class = key.sz;
    best->variant = bestv.variant;
    best->param = bestv.param;
    best->ms = bestv.ms;
  }
  return HIPACC_OK;
}

// ---------------------------------------------------------------------------
// Fused chains
// ---------------------------------------------------------------------------
static bool fused_dtype(int dt) {
  return dt == HIPACC_U8 || dt == HIPACC_I8 || dt == HIPACC_U16 || dt == HIPACC_I16 ||
         dt == HIPACC_I32 || dt == HIPACC_F32;
}

static int prep_fused_local(const hipacc_local_desc* l, int in_dt, int out_dt, FusedLocal& L,
                            int k) {
  if (!l) return set_error(HIPACC_ERR_INVALID_ARG, "fused: stage %d: local is NULL", k);
  if (l->mask_w < 1 || l->mask_h < 1 || l->mask_w > HIPACC_MAX_MASK_DIM ||
      l->mask_h > HIPACC_MAX_MASK_DIM || !(l->mask_w & 1) || !(l->mask_h & 1))
    return set_error(HIPACC_ERR_INVALID_ARG, "fused: stage %d: mask %dx%d (odd, <= %d)", k,
                     l->mask_w, l->mask_h, HIPACC_MAX_MASK_DIM);
  if (l->acc_dtype != HIPACC_F32 && l->acc_dtype != HIPACC_I32)
    return set_error(HIPACC_ERR_INVALID_ARG, "fused: stage %d: acc_dtype must be F32 or I32", k);
  if (l->reduce < HIPACC_REDUCE_SUM || l->reduce > HIPACC_REDUCE_PROD)
    return set_error(HIPACC_ERR_INVALID_ARG, "fused: stage %d: reduce %d", k, l->reduce);
  if (l->boundary < HIPACC_BOUNDARY_UNDEFINED || l->boundary > HIPACC_BOUNDARY_CONSTANT)
    return set_error(HIPACC_ERR_INVALID_ARG, "fused: stage %d: boundary %d", k, l->boundary);
  if (l->cvt != HIPACC_CVT_TRUNC && l->cvt != HIPACC_CVT_ROUND)
    return set_error(HIPACC_ERR_INVALID_ARG, "fused: stage %d: cvt %d", k, l->cvt);
  if (l->use_weights && l->weights_dev)
    return set_error(HIPACC_ERR_UNSUPPORTED, "fused: stage %d: device weights", k);
  const int n = l->mask_w * l->mask_h;
  L.acc_f = l->acc_dtype == HIPACC_F32;
  if (l->use_weights) {
    if (!l->weights_f32 && !l->weights_i32)
      return set_error(HIPACC_ERR_INVALID_ARG, "fused: stage %d: use_weights without weights", k);
    for (int t = 0; t < n; ++t) {
      if (L.acc_f) L.w.f[t] = l->weights_f32 ? l->weights_f32[t] : (float)l->weights_i32[t];
      else L.w.i[t] = l->weights_i32 ? l->weights_i32[t] : (int32_t)l->weights_f32[t];
    }
  }
  bool any = false;
  for (int t = 0; t < n; ++t) {
    L.domain[t] = l->domain ? (l->domain[t] != 0) : 1;
    any = any || L.domain[t];
  }
  if (!any) return set_error(HIPACC_ERR_INVALID_ARG, "fused: stage %d: empty domain", k);
  L.mask_w = l->mask_w;
  L.mask_h = l->mask_h;
  L.reduce = l->reduce;
  L.use_weights = l->use_weights != 0;
  L.fma = l->fp_contract != 0;
  L.boundary = l->boundary;
  L.cvt = l->cvt;
  L.post_div = l->post_div;
  L.in_dtype = in_dt;
  L.out_dtype = out_dt;
  const_views(in_dt, l->const_val, &L.const_f, &L.const_i);
  return HIPACC_OK;
}

static int prep_fused(const hipacc_fused_desc* d, FusedArgs& a) {
  if (!d) return set_error(HIPACC_ERR_INVALID_ARG, "fused: desc is NULL");
  std::memset(&a, 0, sizeof(a));
  int st = check_view(d->in, "in");
  if (st) return st;
  if ((st = check_view(d->out, "out"))) return st;
  const hipacc_view& iv = d->in;
  const hipacc_view& ov = d->out;
  if (!fused_dtype(iv.dtype) || !fused_dtype(ov.dtype))
    return set_error(HIPACC_ERR_UNSUPPORTED, "fused: pixel types %d -> %d", iv.dtype, ov.dtype);
  if (iv.width != ov.width || iv.height != ov.height || batch_of(iv) != batch_of(ov))
    return set_error(HIPACC_ERR_INVALID_ARG, "fused: in (%dx%d) and out (%dx%d) differ in size",
                     iv.width, iv.height, ov.width, ov.height);
  if (d->num_stages < 1 || d->num_stages > HIPACC_MAX_FUSED_STAGES)
    return set_error(HIPACC_ERR_INVALID_ARG, "fused: %d stages (1..%d)", d->num_stages,
                     HIPACC_MAX_FUSED_STAGES);
  int dt = iv.dtype, np = 0, seg = 0;
  a.seg[0][0] = 0;
  for (int k = 0; k < d->num_stages; ++k) {
    const hipacc_fused_stage& g = d->stages[k];
    if (!fused_dtype(g.dtype))
      return set_error(HIPACC_ERR_UNSUPPORTED, "fused: stage %d pixel type %d", k, g.dtype);
    if (g.kind == HIPACC_STAGE_LOCAL) {
      if (a.num_locals == HIPACC_MAX_FUSED_LOCALS)
        return set_error(HIPACC_ERR_UNSUPPORTED, "fused: more than %d local stages",
                         HIPACC_MAX_FUSED_LOCALS);
      if ((st = prep_fused_local(g.local, dt, g.dtype, a.loc[a.num_locals], k))) return st;
      ++a.num_locals;
      a.seg[seg][1] = np;
      a.seg[++seg][0] = np;
    } else if (g.kind == HIPACC_STAGE_POINT) {
      if (g.op < HIPACC_POINT_COPY || g.op > HIPACC_POINT_MAGNITUDE)
        return set_error(HIPACC_ERR_INVALID_ARG, "fused: stage %d: op %d", k, g.op);
      if (g.cvt != HIPACC_CVT_TRUNC && g.cvt != HIPACC_CVT_ROUND)
        return set_error(HIPACC_ERR_INVALID_ARG, "fused: stage %d: cvt %d", k, g.cvt);
      FusedPoint& p = a.pt[np++];
      p.op = g.op;
      p.cvt = g.cvt;
      p.dtype = g.dtype;
      p.alpha = g.alpha;
      p.beta = g.beta;
      const bool binary = g.op >= HIPACC_POINT_ADD;
      if (binary) {
        if ((st = check_view(g.in1, "in1"))) return st;
        if (g.in1.dtype != dt)
          return set_error(HIPACC_ERR_UNSUPPORTED,
                           "fused: stage %d: in1 pixel type %d, the stage's input is %d", k,
                           g.in1.dtype, dt);
        if (g.in1.width != ov.width || g.in1.height != ov.height || batch_of(g.in1) != batch_of(ov))
          return set_error(HIPACC_ERR_INVALID_ARG, "fused: stage %d: in1 window size", k);
        p.in1.data = g.in1.data;
        p.in1.plane = g.in1.plane_stride;
        p.in1.stride = g.in1.stride;
        p.in1.ox = g.in1.offset_x;
        p.in1.oy = g.in1.offset_y;
        p.in1.dtype = g.in1.dtype;
      }
    } else {
      return set_error(HIPACC_ERR_INVALID_ARG, "fused: stage %d: kind %d", k, g.kind);
    }
    dt = g.dtype;
  }
  a.seg[seg][1] = np;
  if (a.num_locals == 0) {
    a.seg[2][0] = a.seg[0][0];
    a.seg[2][1] = a.seg[0][1];
  } else if (a.num_locals == 1) {
    a.seg[2][0] = a.seg[1][0];
    a.seg[2][1] = a.seg[1][1];
    a.seg[1][0] = a.seg[1][1] = 0;
  }
  if (ov.dtype != dt)
    return set_error(HIPACC_ERR_INVALID_ARG, "fused: out pixel type %d, the last stage writes %d",
                     ov.dtype, dt);
  a.in = iv.data;
  a.out = ov.data;
  a.in_plane = iv.plane_stride;
  a.out_plane = ov.plane_stride;
  a.in_stride = iv.stride;
  a.out_stride = ov.stride;
  a.in_dtype = iv.dtype;
  a.out_dtype = ov.dtype;
  a.in_imgw = iv.img_width;
  a.in_imgh = iv.img_height;
  a.in_ox = iv.offset_x;
  a.in_oy = iv.offset_y;
  a.out_ox = ov.offset_x;
  a.out_oy = ov.offset_y;
  a.w = ov.width;
  a.h = ov.height;
  a.batch = batch_of(ov);
  return HIPACC_OK;
}

}  // namespace hh

extern "C" {

int hipacc_hip_local(const hipacc_local_desc* d, void* stream) {
  return run_local(d, kVarAuto, 0, static_cast<hipStream_t>(stream));
}

int hipacc_hip_local_variant(const hipacc_local_desc* d, int32_t variant, int32_t param, void* stream) {
  if (variant < kVarAuto || variant > kVarRing)
    return set_error(HIPACC_ERR_INVALID_ARG, "local_variant: variant %d", variant);
  return run_local(d, variant, param, static_cast<hipStream_t>(stream));
}

int hipacc_hip_separable(const hipacc_separable_desc* d, void* stream) {
  return run_separable(d, kVarAuto, 0, static_cast<hipStream_t>(stream));
}

int hipacc_hip_separable_variant(const hipacc_separable_desc* d, int32_t variant, int32_t param,
                                 void* stream) {
  if (variant < kVarAuto || variant > kVarRing)
    return set_error(HIPACC_ERR_INVALID_ARG, "separable_variant: variant %d", variant);
  return run_separable(d, variant, param, static_cast<hipStream_t>(stream));
}

int hipacc_hip_tune_local(const hipacc_local_desc* d, int32_t iters, hipacc_tune_entry* best,
                          void* stream) {
  LocalArgs a;
  SepArgs sa;
  int accf = 0;
  bool box = false;
  int st = prep_local(d, a, accf, box, sa);
  if (st) return st;
  hipStream_t s = static_cast<hipStream_t>(stream);
  TuneRotation rot;
  if ((st = rot.init(d->in, d->out))) return st;
  const int sz = tune_size_class(a.g.is_w, a.g.is_h, a.g.batch);
  if (box) {
    TuneKey k{kTuneBox, sa.kx, sa.ky, sa.g.in_dtype, sa.g.out_dtype, 0, sa.g.boundary, 0, 0, sz};
    return tune_candidates(
        [&](int v, int p, int r) {
          SepArgs b = sa;
          rot.apply(b.g, r);
          return launch_separable(b, 0, 0, s, v, p);
        },
        k, iters, s, best, rot.n);
  }
  TuneKey k{kTuneLocal, a.mask_w, a.mask_h, a.g.in_dtype, a.g.out_dtype, accf, a.g.boundary,
            a.fma, a.wdev != nullptr, sz};
  return tune_candidates(
      [&](int v, int p, int r) {
        LocalArgs b = a;
        rot.apply(b.g, r);
        return launch_local(b, accf, s, v, p);
      },
      k, iters, s, best, rot.n);
}

int hipacc_hip_tune_separable(const hipacc_separable_desc* d, int32_t iters, hipacc_tune_entry* best,
                              void* stream) {
  SepArgs a;
  int accf = 0;
  int st = prep_separable(d, a, accf);
  if (st) return st;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int fma = d->fp_contract != 0;
  TuneRotation rot;
  if ((st = rot.init(d->in, d->out))) return st;
  TuneKey k{a.box ? kTuneBox : kTuneSeparable, a.kx, a.ky, a.g.in_dtype, a.g.out_dtype, accf,
            a.g.boundary, fma, 0, tune_size_class(a.g.is_w, a.g.is_h, a.g.batch)};
  return tune_candidates(
      [&](int v, int p, int r) {
        SepArgs b = a;
        rot.apply(b.g, r);
        return launch_separable(b, accf, fma, s, v, p);
      },
      k, iters, s, best, rot.n);
}

int hipacc_hip_tune_set(const hipacc_tune_entry* e) {
  if (!e) return set_error(HIPACC_ERR_INVALID_ARG, "tune_set: NULL entry");
  if (e->variant < kVarTile || e->variant > kVarRing)
    return set_error(HIPACC_ERR_INVALID_ARG, "tune_set: variant %d", e->variant);
  TuneKey k{e->kind, e->kx, e->ky, e->in_dtype, e->out_dtype, e->acc_dtype == HIPACC_F32,
            e->boundary, e->fp_contract, e->weights_dev, e->size_class};
  tune_store(k, TuneVal{e->variant, e->param, e->ms});
  return HIPACC_OK;
}

int hipacc_hip_tune_save(const char* path) {
  if (!path) return set_error(HIPACC_ERR_INVALID_ARG, "tune_save: NULL path");
  return tune_save(path);
}

int hipacc_hip_tune_load(const char* path) {
  if (!path) return set_error(HIPACC_ERR_INVALID_ARG, "tune_load: NULL path");
  return tune_load(path);
}

int hipacc_hip_point(const hipacc_point_desc* d0, void* stream) {
  if (!d0) return set_error(HIPACC_ERR_INVALID_ARG, "point: desc is NULL");
  // Vector pixels: every operator here is elementwise, so a 4-channel image
  // is the scalar image of its interleaved channels (4x the columns).
  hipacc_point_desc dv = *d0;
  const bool bin0 = d0->op >= HIPACC_POINT_ADD && d0->op <= HIPACC_POINT_MAGNITUDE;
  const bool vin = dtype_is_vec(d0->in0.dtype), vout = dtype_is_vec(d0->out.dtype);
  if (vin || vout || (bin0 && dtype_is_vec(d0->in1.dtype))) {
    if (!vin || !vout || (bin0 && !dtype_is_vec(d0->in1.dtype)))
      return set_error(HIPACC_ERR_INVALID_ARG, "point: vector and scalar pixels mixed");
    dv.in0 = channel_view(d0->in0);
    dv.out = channel_view(d0->out);
    if (bin0) dv.in1 = channel_view(d0->in1);
  }
  const hipacc_point_desc* d = &dv;
  int st = check_view(d->in0, "in0");
  if (st) return st;
  if ((st = check_view(d->out, "out"))) return st;
  const bool binary = d->op >= HIPACC_POINT_ADD && d->op <= HIPACC_POINT_MAGNITUDE;
  if (d->op < HIPACC_POINT_COPY || d->op > HIPACC_POINT_MAGNITUDE)
    return set_error(HIPACC_ERR_INVALID_ARG, "point: op %d", d->op);
  if (binary && (st = check_view(d->in1, "in1"))) return st;
  const hipacc_view& o = d->out;
  if (d->in0.width < o.width || d->in0.height < o.height ||
      (binary && (d->in1.width < o.width || d->in1.height < o.height)))
    return set_error(HIPACC_ERR_INVALID_ARG, "point: input window smaller than iteration space");
  if (batch_of(d->in0) != batch_of(o) || (binary && batch_of(d->in1) != batch_of(o)))
    return set_error(HIPACC_ERR_INVALID_ARG, "point: batch mismatch");
  if (d->cvt != HIPACC_CVT_TRUNC && d->cvt != HIPACC_CVT_ROUND)
    return set_error(HIPACC_ERR_INVALID_ARG, "point: cvt %d", d->cvt);
  PointArgs a;
  std::memset(&a, 0, sizeof(a));
  a.in0 = d->in0.data;
  a.in1 = binary ? d->in1.data : nullptr;
  a.out = o.data;
  a.plane0 = d->in0.plane_stride;
  a.plane1 = binary ? d->in1.plane_stride : 0;
  a.plane_out = o.plane_stride;
  a.stride0 = d->in0.stride;
  a.stride1 = binary ? d->in1.stride : 0;
  a.stride_out = o.stride;
  a.ox0 = d->in0.offset_x;
  a.oy0 = d->in0.offset_y;
  a.ox1 = binary ? d->in1.offset_x : 0;
  a.oy1 = binary ? d->in1.offset_y : 0;
  a.oxo = o.offset_x;
  a.oyo = o.offset_y;
  a.w = o.width;
  a.h = o.height;
  a.batch = batch_of(o);
  a.op = d->op;
  a.cvt = d->cvt;
  a.alpha = d->alpha;
  a.beta = d->beta;
  const int e0 = dtype_bytes(d->in0.dtype), eo = dtype_bytes(o.dtype);
  a.vec_ok = rows_aligned(d->in0, 4 * e0) && rows_aligned(o, 4 * eo) &&
             (!binary || rows_aligned(d->in1, 4 * dtype_bytes(d->in1.dtype)));
  hipStream_t s = static_cast<hipStream_t>(stream);
  Timer tm(s);
  return tm.finish(launch_point(a, d->in0.dtype, binary ? d->in1.dtype : d->in0.dtype, o.dtype, s));
}

int hipacc_hip_fused(const hipacc_fused_desc* d, void* stream) {
  FusedArgs a;
  const int st = prep_fused(d, a);
  if (st) return st;
  hipStream_t s = static_cast<hipStream_t>(stream);
  Timer tm(s);
  return tm.finish(launch_fused(a, s));
}

size_t hipacc_hip_reduce_workspace_bytes(void) { return kRedWorkspace; }

int hipacc_hip_reduce(const hipacc_view* img, int32_t reduce, int32_t acc_dtype, void* result_dev,
                      void* workspace, void* stream) {
  if (!img || !result_dev) return set_error(HIPACC_ERR_INVALID_ARG, "reduce: NULL argument");
  int st = check_view(*img, "img");
  if (st) return st;
  if (batch_of(*img) != 1) return set_error(HIPACC_ERR_UNSUPPORTED, "reduce: batch > 1");
  if (dtype_is_vec(img->dtype)) return set_error(HIPACC_ERR_UNSUPPORTED, "reduce: vector pixels");
  if (reduce < HIPACC_REDUCE_SUM || reduce > HIPACC_REDUCE_PROD)
    return set_error(HIPACC_ERR_INVALID_ARG, "reduce: mode %d", reduce);
  if (acc_dtype != HIPACC_F32 && acc_dtype != HIPACC_U32 && acc_dtype != HIPACC_I32)
    return set_error(HIPACC_ERR_INVALID_ARG, "reduce: acc_dtype %d", acc_dtype);
  if ((img->dtype == HIPACC_F32) != (acc_dtype == HIPACC_F32))
    return set_error(HIPACC_ERR_UNSUPPORTED, "reduce: float images need a float accumulator");
  if (!workspace) {
    int dev;
    if ((st = current_device(&dev))) return st;
    if ((st = ensure_workspace(dev, &workspace))) return st;
  }
  ReduceArgs a;
  a.in = img->data;
  a.stride = img->stride;
  a.ox = img->offset_x;
  a.oy = img->offset_y;
  a.w = img->width;
  a.h = img->height;
  a.mode = reduce;
  a.vec_ok = rows_aligned(*img, 16);
  a.result = result_dev;
  a.partials = static_cast<unsigned long long*>(workspace);
  hipStream_t s = static_cast<hipStream_t>(stream);
  Timer tm(s);
  return tm.finish(launch_reduce(a, img->dtype, acc_dtype, s));
}

int hipacc_hip_reduce_sync(const hipacc_view* img, int32_t reduce, int32_t acc_dtype,
                           void* result_host, void* stream) {
  if (!result_host) return set_error(HIPACC_ERR_INVALID_ARG, "reduce_sync: result_host is NULL");
  int dev, st;
  if ((st = current_device(&dev))) return st;
  void* ws = nullptr;
  if ((st = ensure_workspace(dev, &ws))) return st;
  // The scalar result lives at the end of the library workspace.  Callers on
  // other host threads share it and its partials, so the device lock is held
  // from the launch until the result is on the host.
  void* res = static_cast<char*>(ws) + 1024 * sizeof(unsigned long long) + 64;
  hipStream_t s = static_cast<hipStream_t>(stream);
  std::lock_guard<std::mutex> lk(g_dev[dev].mu);
  if ((st = hipacc_hip_reduce(img, reduce, acc_dtype, res, ws, stream))) return st;
  hipError_t e = hipMemcpyAsync(result_host, res, 4, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  return hip_status(e, "reduce_sync copy-back");
}

int hipacc_hip_histogram(const hipacc_view* img, int32_t num_bins, float scale, uint32_t* bins_dev,
                         void* stream) {
  if (!img || !bins_dev) return set_error(HIPACC_ERR_INVALID_ARG, "histogram: NULL argument");
  int st = check_view(*img, "img");
  if (st) return st;
  if (batch_of(*img) != 1) return set_error(HIPACC_ERR_UNSUPPORTED, "histogram: batch > 1");
  if (num_bins < 1 || num_bins > HIPACC_MAX_BINS)
    return set_error(HIPACC_ERR_INVALID_ARG, "histogram: num_bins %d", num_bins);
  HistArgs a;
  a.in = img->data;
  a.stride = img->stride;
  a.ox = img->offset_x;
  a.oy = img->offset_y;
  a.w = img->width;
  a.h = img->height;
  a.nbins = num_bins;
  a.scale = scale;
  a.vec_ok = rows_aligned(*img, 16);
  a.bins = bins_dev;
  hipStream_t s = static_cast<hipStream_t>(stream);
  void* ws = nullptr;
  if ((st = stream_workspace(s, &ws))) return st;
  a.partials = static_cast<uint32_t*>(ws);
  Timer tm(s);
  return tm.finish(launch_histogram(a, img->dtype, s));
}

int hipacc_hip_bilateral(const hipacc_bilateral_desc* d, void* stream) {
  if (!d) return set_error(HIPACC_ERR_INVALID_ARG, "bilateral: desc is NULL");
  BilatArgs a;
  std::memset(&a, 0, sizeof(a));
  int st = make_geom(a.g, d->in, d->out, d->boundary, d->const_val, d->cvt, 0.0f);
  if (st) return st;
  if (!((d->in.dtype == HIPACC_F32 && d->out.dtype == HIPACC_F32) ||
        (d->in.dtype == HIPACC_U8 && d->out.dtype == HIPACC_U8)))
    return set_error(HIPACC_ERR_UNSUPPORTED, "bilateral: F32->F32 or U8->U8 only");
  if (d->ksize < 1 || d->ksize > HIPACC_MAX_MASK_DIM || !(d->ksize & 1))
    return set_error(HIPACC_ERR_INVALID_ARG, "bilateral: ksize %d", d->ksize);
  if (!(d->sigma_r > 0.0f)) return set_error(HIPACC_ERR_INVALID_ARG, "bilateral: sigma_r must be > 0");
  if (!d->spatial && !(d->sigma_d > 0.0f))
    return set_error(HIPACC_ERR_INVALID_ARG, "bilateral: sigma_d must be > 0");
  a.ksize = d->ksize;
  const int r = d->ksize / 2;
  for (int dy = 0; dy < d->ksize; ++dy)
    for (int dx = 0; dx < d->ksize; ++dx) {
      const int t = dy * d->ksize + dx;
      float w;
      if (d->spatial) {
        w = d->spatial[t];
      } else {
        const double ddx = dx - r, ddy = dy - r;
        w = (float)std::exp(-(ddx * ddx + ddy * ddy) / (2.0 * (double)d->sigma_d * d->sigma_d));
      }
      a.lsp[t] = w > 0.0f ? std::log2(w) : -INFINITY;
    }
  const float c_r = 0.5f / (d->sigma_r * d->sigma_r);
  a.k2 = -c_r * 1.44269504088896340736f;
  a.in_vec_ok = d->in.dtype == HIPACC_F32 && rows_aligned(d->in, 16) && (d->in.offset_x % 4) == 0;
  a.out_vec_ok = rows_aligned(d->out, 4 * dtype_bytes(d->out.dtype));
  hipStream_t s = static_cast<hipStream_t>(stream);
  Timer tm(s);
  return tm.finish(launch_bilateral(a, s));
}

}  // extern "C"

namespace hh {

static size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

// Workspace layout of the convolution adjoint: [flipped mask][ring][dw partials].
struct ConvGradWs {
  size_t flip, ring, part, total;
};
static ConvGradWs conv_grad_ws(const hipacc_local_grad_desc* d) {
  ConvGradWs w;
  const int kk = d->mask_w * d->mask_h;
  const bool dx = d->grad_in.data != nullptr, dw = d->grad_weights_dev != nullptr;
  const int nb = d->in.batch > 1 ? d->in.batch : 1;
  w.flip = 0;
  w.ring = align256(sizeof(float) * kMaxTaps);
  const size_t ring = (dx && d->boundary != HIPACC_BOUNDARY_CONSTANT)
                          ? sizeof(float) * (size_t)nb *
                                (size_t)ring_size(d->in.width, d->in.height, d->mask_w / 2,
                                                  d->mask_h / 2)
                          : 0;
  w.part = w.ring + align256(ring);
  w.total = w.part + (dw ? align256(sizeof(float) * (size_t)kGradBlocks * kk) : 0);
  return w;
}

static int check_local_grad(const hipacc_local_grad_desc* d) {
  if (!d) return set_error(HIPACC_ERR_INVALID_ARG, "local_grad: desc is NULL");
  int st = grad_image(d->in, "in");
  if (st) return st;
  if ((st = grad_image(d->grad_out, "grad_out"))) return st;
  if (!same_shape(d->in, d->grad_out))
    return set_error(HIPACC_ERR_INVALID_ARG, "local_grad: grad_out differs from in in size");
  if (d->grad_in.data) {
    if ((st = grad_image(d->grad_in, "grad_in"))) return st;
    if (!same_shape(d->in, d->grad_in))
      return set_error(HIPACC_ERR_INVALID_ARG, "local_grad: grad_in differs from in in size");
  }
  if (d->mask_w < 1 || d->mask_h < 1 || d->mask_w > HIPACC_MAX_MASK_DIM ||
      d->mask_h > HIPACC_MAX_MASK_DIM || !(d->mask_w & 1) || !(d->mask_h & 1))
    return set_error(HIPACC_ERR_INVALID_ARG, "local_grad: mask %dx%d (odd, <= %d)", d->mask_w,
                     d->mask_h, HIPACC_MAX_MASK_DIM);
  if (!d->weights_dev && !d->weights_f32)
    return set_error(HIPACC_ERR_INVALID_ARG, "local_grad: no weights");
  if (d->boundary < HIPACC_BOUNDARY_UNDEFINED || d->boundary > HIPACC_BOUNDARY_CONSTANT)
    return set_error(HIPACC_ERR_INVALID_ARG, "local_grad: boundary %d", d->boundary);
  return HIPACC_OK;
}

static int run_local_grad(const hipacc_local_grad_desc* d, int variant, char* ws, bool timed,
                          hipStream_t s);

}  // namespace hh

extern "C" {

size_t hipacc_hip_local_grad_workspace_bytes(const hipacc_local_grad_desc* d) {
  if (check_local_grad(d)) return 0;
  return conv_grad_ws(d).total;
}

int hipacc_hip_local_grad(const hipacc_local_grad_desc* d, void* stream) {
  return hipacc_hip_local_grad_variant(d, 0, stream);
}

int hipacc_hip_local_grad_variant(const hipacc_local_grad_desc* d, int32_t variant,
                                  void* stream) {
  int st = check_local_grad(d);
  if (st) return st;
  if (variant != kVarAuto && variant != kVarTile && variant != kVarStream)
    return set_error(HIPACC_ERR_INVALID_ARG, "local_grad: variant %d", variant);
  if (!d->grad_in.data && !d->grad_weights_dev) return HIPACC_OK;
  char* ws = static_cast<char*>(d->workspace);
  if (!ws) {
    void* p = nullptr;
    if ((st = ensure_grad_workspace(conv_grad_ws(d).total, &p))) return st;
    ws = static_cast<char*>(p);
  }
  return run_local_grad(d, variant, ws, true, static_cast<hipStream_t>(stream));
}

int hipacc_hip_local_grad_noalloc(const hipacc_local_grad_desc* d, void* stream) {
  int st = check_local_grad(d);
  if (st) return st;
  if (d->workspace)
    return set_error(HIPACC_ERR_INVALID_ARG,
                     "local_grad_noalloc: desc->workspace must be NULL (the library's static "
                     "scratch is the workspace)");
  if (!d->grad_in.data && !d->grad_weights_dev) return HIPACC_OK;
  const size_t need = conv_grad_ws(d).total;
  if (need > (size_t)HIPACC_NOALLOC_SCRATCH_BYTES)
    return set_error(HIPACC_ERR_INVALID_ARG,
                     "local_grad_noalloc: needs %zu bytes of workspace, more than the %d-byte "
                     "static scratch (use hipacc_hip_local_grad)",
                     need, (int)HIPACC_NOALLOC_SCRATCH_BYTES);
  int dev = 0;
  if ((st = current_device(&dev))) return st;
  void* p = nullptr;
  if ((st = scratch_address(dev, &p))) return st;
  return run_local_grad(d, kVarAuto, static_cast<char*>(p), false,
                        static_cast<hipStream_t>(stream));
}

}  // extern "C"

namespace hh {

// The adjoint's launches with workspace `ws` (conv_grad_ws(d).total bytes;
// nothing in it needs to start at zero: every word is written before it is
// read).  `timed`: honour hipacc_hip_set_timing.
static int run_local_grad(const hipacc_local_grad_desc* d, int variant, char* ws, bool timed,
                          hipStream_t s) {
  int st = HIPACC_OK;
  const bool dx = d->grad_in.data != nullptr, dw = d->grad_weights_dev != nullptr;
  const int kx = d->mask_w, ky = d->mask_h, kk = kx * ky;
  const ConvGradWs lay = conv_grad_ws(d);
  ConvGradArgs a;
  std::memset(&a, 0, sizeof(a));
  a.x = static_cast<const float*>(d->in.data);
  a.gy = static_cast<const float*>(d->grad_out.data);
  a.gx = dx ? static_cast<float*>(d->grad_in.data) : nullptr;
  a.x_plane = d->in.plane_stride;
  a.gy_plane = d->grad_out.plane_stride;
  a.gx_plane = dx ? d->grad_in.plane_stride : 0;
  a.x_stride = d->in.stride;
  a.gy_stride = d->grad_out.stride;
  a.gx_stride = dx ? d->grad_in.stride : 0;
  a.w = d->in.width;
  a.h = d->in.height;
  a.batch = batch_of(d->in);
  a.kx = kx;
  a.ky = ky;
  a.boundary = d->boundary;
  a.cval = (float)d->const_val;
  a.vec_ok = rows_aligned(d->in, 16) && rows_aligned(d->grad_out, 16);
  a.wdev = d->weights_dev;
  if (!a.wdev) std::memcpy(a.wh, d->weights_f32, sizeof(float) * kk);
  a.partials = reinterpret_cast<float*>(ws + lay.part);
  a.dw_variant = variant;
  Timer tm(s, timed);
  // dw's partials first: their reduction then rides on the fold's launch
  PartialSum sum{};
  if (dw && (st = launch_conv_dw(a, d->grad_weights_dev, s, &sum))) return tm.finish(st);
  bool summed = !dw;
  if (dx) {
    // E over the image: the forward operator on dY with the flipped mask and
    // a zero border.
    hipacc_local_desc f;
    std::memset(&f, 0, sizeof(f));
    f.in = d->grad_out;
    f.out = d->grad_in;
    f.mask_w = kx;
    f.mask_h = ky;
    float flipped[kMaxTaps];
    const bool folds = d->boundary != HIPACC_BOUNDARY_CONSTANT;
    float* ring = reinterpret_cast<float*>(ws + lay.ring);
    float* wf = d->weights_dev ? reinterpret_cast<float*>(ws + lay.flip) : nullptr;
    // the ring pass (which also writes the flipped device mask) before the
    // image pass that reads that mask
    if (folds) {
      if ((st = launch_conv_ring(a, ring, wf, s))) return tm.finish(st);
    } else if (wf) {
      if ((st = launch_flip(d->weights_dev, wf, kk, s))) return tm.finish(st);
    }
    if (d->weights_dev) {
      f.weights_dev = wf;
    } else {
      for (int t = 0; t < kk; ++t) flipped[t] = d->weights_f32[kk - 1 - t];
      f.weights_f32 = flipped;
    }
    f.use_weights = 1;
    f.acc_dtype = HIPACC_F32;
    f.reduce = HIPACC_REDUCE_SUM;
    f.boundary = HIPACC_BOUNDARY_CONSTANT;
    f.const_val = 0.0;
    f.cvt = HIPACC_CVT_TRUNC;
    f.fp_contract = 1;
    LocalArgs la;
    SepArgs sa;
    int accf = 0;
    bool box = false;
    if ((st = prep_local(&f, la, accf, box, sa))) return tm.finish(st);
    if ((st = launch_local(la, accf, s, kVarAuto, 0))) return tm.finish(st);
    if (folds) {
      if ((st = launch_fold_add(a.gx, a.gx_plane, a.gx_stride, a.w, a.h, a.batch, kx / 2, ky / 2,
                                d->boundary, ring, s, dw ? &sum : nullptr)))
        return tm.finish(st);
      summed = true;
    }
  }
  if (!summed) st = launch_partial_sum(sum, s);
  return tm.finish(st);
}

}  // namespace hh

extern "C" {

// Spatial table and argument checks shared by the bilateral entry points.
static int bilat_common(const hipacc_bilateral_grad_desc* d, float* sp) {
  if (d->ksize < 1 || d->ksize > HIPACC_MAX_MASK_DIM || !(d->ksize & 1))
    return set_error(HIPACC_ERR_INVALID_ARG, "bilateral: ksize %d", d->ksize);
  if (!(d->sigma_r > 0.0f)) return set_error(HIPACC_ERR_INVALID_ARG, "bilateral: sigma_r must be > 0");
  if (!d->spatial && !(d->sigma_d > 0.0f))
    return set_error(HIPACC_ERR_INVALID_ARG, "bilateral: sigma_d must be > 0");
  if (d->boundary < HIPACC_BOUNDARY_UNDEFINED || d->boundary > HIPACC_BOUNDARY_CONSTANT)
    return set_error(HIPACC_ERR_INVALID_ARG, "bilateral: boundary %d", d->boundary);
  const int r = d->ksize / 2;
  for (int dy = 0; dy < d->ksize; ++dy)
    for (int dx = 0; dx < d->ksize; ++dx) {
      const int t = dy * d->ksize + dx;
      if (d->spatial) {
        sp[t] = d->spatial[t];
      } else {
        const double ddx = dx - r, ddy = dy - r;
        sp[t] = (float)std::exp(-(ddx * ddx + ddy * ddy) / (2.0 * (double)d->sigma_d * d->sigma_d));
      }
    }
  return HIPACC_OK;
}

int hipacc_hip_bilateral_train(const hipacc_bilateral_grad_desc* d, void* stream) {
  if (!d) return set_error(HIPACC_ERR_INVALID_ARG, "bilateral_train: desc is NULL");
  int st;
  if ((st = grad_image(d->in, "in")) || (st = grad_image(d->out, "out")) ||
      (st = grad_image(d->norm, "norm")))
    return st;
  if (!same_shape(d->in, d->out) || !same_shape(d->in, d->norm))
    return set_error(HIPACC_ERR_INVALID_ARG, "bilateral_train: in, out and norm differ in size");
  float sp[kMaxTaps];
  if ((st = bilat_common(d, sp))) return st;
  BilatArgs a;
  std::memset(&a, 0, sizeof(a));
  if ((st = make_geom(a.g, d->in, d->out, d->boundary, d->const_val, HIPACC_CVT_TRUNC, 0.0f)))
    return st;
  a.ksize = d->ksize;
  for (int t = 0; t < d->ksize * d->ksize; ++t) a.lsp[t] = sp[t] > 0.0f ? std::log2(sp[t]) : -INFINITY;
  const float c_r = 0.5f / (d->sigma_r * d->sigma_r);
  a.k2 = -c_r * 1.44269504088896340736f;
  a.in_vec_ok = rows_aligned(d->in, 16);
  a.out_vec_ok = rows_aligned(d->out, 16);
  a.norm = static_cast<float*>(d->norm.data);
  a.norm_plane = d->norm.plane_stride;
  a.norm_stride = d->norm.stride;
  hipStream_t s = static_cast<hipStream_t>(stream);
  Timer tm(s);
  return tm.finish(launch_bilateral(a, s));
}

size_t hipacc_hip_bilateral_grad_workspace_bytes(const hipacc_bilateral_grad_desc* d) {
  if (!d) return 0;
  const int nb = d->in.batch > 1 ? d->in.batch : 1;
  const int r = d->ksize / 2;
  const size_t ring = (d->grad_in.data && d->boundary != HIPACC_BOUNDARY_CONSTANT)
                          ? sizeof(float) * (size_t)nb *
                                (size_t)ring_size(d->in.width, d->in.height, r, r)
                          : 0;
  return align256(ring) + align256(sizeof(float) * kGradBlocks);
}

int hipacc_hip_bilateral_grad(const hipacc_bilateral_grad_desc* d, void* stream) {
  if (!d) return set_error(HIPACC_ERR_INVALID_ARG, "bilateral_grad: desc is NULL");
  const bool dx = d->grad_in.data != nullptr, ds = d->grad_sigma_r_dev != nullptr;
  int st;
  if ((st = grad_image(d->in, "in")) || (st = grad_image(d->out, "out")) ||
      (st = grad_image(d->norm, "norm")) || (st = grad_image(d->grad_out, "grad_out")))
    return st;
  if (dx && (st = grad_image(d->grad_in, "grad_in"))) return st;
  if (!same_shape(d->in, d->out) || !same_shape(d->in, d->norm) ||
      !same_shape(d->in, d->grad_out) || (dx && !same_shape(d->in, d->grad_in)))
    return set_error(HIPACC_ERR_INVALID_ARG, "bilateral_grad: images differ in size");
  BilatGradArgs a;
  std::memset(&a, 0, sizeof(a));
  if ((st = bilat_common(d, a.sp))) return st;
  if (!dx && !ds) return HIPACC_OK;
  const int kk = d->ksize * d->ksize;
  a.sym = 1;
  for (int t = 0; t < kk; ++t) {
    a.lsp[t] = a.sp[t] > 0.0f ? std::log2(a.sp[t]) : -INFINITY;
    a.sym = a.sym && a.sp[t] == a.sp[kk - 1 - t];
  }
  a.x = static_cast<const float*>(d->in.data);
  a.y = static_cast<const float*>(d->out.data);
  a.d = static_cast<const float*>(d->norm.data);
  a.gy = static_cast<const float*>(d->grad_out.data);
  a.gx = dx ? static_cast<float*>(d->grad_in.data) : nullptr;
  a.x_plane = d->in.plane_stride;
  a.y_plane = d->out.plane_stride;
  a.d_plane = d->norm.plane_stride;
  a.gy_plane = d->grad_out.plane_stride;
  a.gx_plane = dx ? d->grad_in.plane_stride : 0;
  a.x_stride = d->in.stride;
  a.y_stride = d->out.stride;
  a.d_stride = d->norm.stride;
  a.gy_stride = d->grad_out.stride;
  a.gx_stride = dx ? d->grad_in.stride : 0;
  a.w = d->in.width;
  a.h = d->in.height;
  a.batch = batch_of(d->in);
  a.ksize = d->ksize;
  a.boundary = d->boundary;
  a.cval = (float)d->const_val;
  a.c = 0.5f / (d->sigma_r * d->sigma_r);
  a.k2 = -a.c * 1.44269504088896340736f;
  a.vec_ok = rows_aligned(d->in, 16) && rows_aligned(d->out, 16) && rows_aligned(d->norm, 16) &&
             rows_aligned(d->grad_out, 16) && (!dx || rows_aligned(d->grad_in, 16));
  const size_t total = hipacc_hip_bilateral_grad_workspace_bytes(d);
  char* ws = static_cast<char*>(d->workspace);
  if (!ws) {
    void* p = nullptr;
    if ((st = ensure_grad_workspace(total, &p))) return st;
    ws = static_cast<char*>(p);
  }
  const size_t part = total - align256(sizeof(float) * kGradBlocks);
  a.partials = reinterpret_cast<float*>(ws + part);
  float* ring = (dx && d->boundary != HIPACC_BOUNDARY_CONSTANT) ? reinterpret_cast<float*>(ws)
                                                                : nullptr;
  hipStream_t s = static_cast<hipStream_t>(stream);
  Timer tm(s);
  return tm.finish(
      launch_bilateral_grad(a, ds ? d->grad_sigma_r_dev : nullptr, d->sigma_r, ring, s));
}

int hipacc_hip_resample(const hipacc_resample_desc* d, void* stream) {
  if (!d) return set_error(HIPACC_ERR_INVALID_ARG, "resample: desc is NULL");
  int st = check_view(d->in, "in");
  if (st) return st;
  if ((st = check_view(d->out, "out"))) return st;
  if ((d->in.dtype != HIPACC_F32 && d->in.dtype != HIPACC_U8) || d->out.dtype != d->in.dtype)
    return set_error(HIPACC_ERR_UNSUPPORTED,
                     "resample: F32 or U8 images, the output of the input's type");
  if (batch_of(d->in) != 1 || batch_of(d->out) != 1)
    return set_error(HIPACC_ERR_UNSUPPORTED, "resample: single images only");
  if (d->interp < HIPACC_INTERP_NN || d->interp > HIPACC_INTERP_L3)
    return set_error(HIPACC_ERR_INVALID_ARG, "resample: interp %d", d->interp);
  if (d->op < HIPACC_RESAMPLE_COPY || d->op > HIPACC_RESAMPLE_ADD)
    return set_error(HIPACC_ERR_INVALID_ARG, "resample: op %d", d->op);
  if (d->boundary < HIPACC_BOUNDARY_UNDEFINED || d->boundary > HIPACC_BOUNDARY_CONSTANT)
    return set_error(HIPACC_ERR_INVALID_ARG, "resample: boundary %d", d->boundary);
  if (d->op != HIPACC_RESAMPLE_COPY) {
    if ((st = check_view(d->in2, "in2"))) return st;
    if (d->in2.dtype != HIPACC_F32 || d->out.dtype != HIPACC_F32 ||
        d->in2.width != d->out.width || d->in2.height != d->out.height || batch_of(d->in2) != 1)
      return set_error(HIPACC_ERR_INVALID_ARG,
                       "resample: SUB / ADD take F32 images, in2 of out's size");
  }
  ResampleArgs a;
  std::memset(&a, 0, sizeof(a));
  a.in = d->in.data;
  a.in2 = static_cast<const float*>(d->in2.data);
  a.out = d->out.data;
  a.in_dtype = d->in.dtype;
  a.out_dtype = d->out.dtype;
  a.xf = d->xf;
  a.yf = d->yf;
  a.in_stride = d->in.stride;
  a.in2_stride = d->in2.stride;
  a.out_stride = d->out.stride;
  a.in_imgw = d->in.img_width;
  a.in_imgh = d->in.img_height;
  a.lo_x = d->in.offset_x;
  a.lo_y = d->in.offset_y;
  a.hi_x = d->in.offset_x + d->in.width;
  a.hi_y = d->in.offset_y + d->in.height;
  a.in2_ox = d->in2.offset_x;
  a.in2_oy = d->in2.offset_y;
  a.out_ox = d->out.offset_x;
  a.out_oy = d->out.offset_y;
  a.w = d->out.width;
  a.h = d->out.height;
  a.sx = (float)d->in.width / (float)d->out.width;
  a.sy = (float)d->in.height / (float)d->out.height;
  a.interp = d->interp;
  a.op = d->op;
  a.boundary = d->boundary;
  {
    float cf;
    int ci;
    const_views(d->in.dtype, d->const_val, &cf, &ci);
    a.cval = cf;   // the constant in the pixel type, seen as float
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  Timer tm(s);
  return tm.finish(launch_resample(a, s));
}

int hipacc_hip_alloc2d(void** ptr, int32_t* stride_elems, int32_t width, int32_t height,
                       int32_t elem_bytes, int32_t alignment) {
  if (!ptr || !stride_elems || width <= 0 || height <= 0 ||
      !(elem_bytes == 1 || elem_bytes == 2 || elem_bytes == 4 || elem_bytes == 8 ||
        elem_bytes == 16))
    return set_error(HIPACC_ERR_INVALID_ARG, "alloc2d: bad arguments");
  if (alignment <= 0) alignment = 256;
  if (alignment % elem_bytes) return set_error(HIPACC_ERR_INVALID_ARG, "alloc2d: alignment");
  const int per = alignment / elem_bytes;
  const long long stride = ((long long)width + per - 1) / per * per;
  *stride_elems = (int32_t)stride;
  return hip_status(hipMalloc(ptr, (size_t)stride * height * elem_bytes), "hipMalloc");
}

int hipacc_hip_free(void* ptr) { return hip_status(hipFree(ptr), "hipFree"); }

int hipacc_hip_upload2d(void* dst_dev, int32_t dst_stride_elems, const void* src_host, int32_t width,
                        int32_t height, int32_t elem_bytes, void* stream) {
  if (!dst_dev || !src_host) return set_error(HIPACC_ERR_INVALID_ARG, "upload2d: NULL pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipError_t e = hipMemcpy2DAsync(dst_dev, (size_t)dst_stride_elems * elem_bytes, src_host,
                                  (size_t)width * elem_bytes, (size_t)width * elem_bytes, height,
                                  hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  return hip_status(e, "upload2d");
}

int hipacc_hip_download2d(void* dst_host, const void* src_dev, int32_t src_stride_elems, int32_t width,
                          int32_t height, int32_t elem_bytes, void* stream) {
  if (!dst_host || !src_dev) return set_error(HIPACC_ERR_INVALID_ARG, "download2d: NULL pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipError_t e = hipMemcpy2DAsync(dst_host, (size_t)width * elem_bytes, src_dev,
                                  (size_t)src_stride_elems * elem_bytes, (size_t)width * elem_bytes,
                                  height, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  return hip_status(e, "download2d");
}

int hipacc_hip_copy_region(const hipacc_view* src, const hipacc_view* dst, void* stream) {
  if (!src || !dst) return set_error(HIPACC_ERR_INVALID_ARG, "copy_region: NULL argument");
  int st = check_view(*src, "src");
  if (st) return st;
  if ((st = check_view(*dst, "dst"))) return st;
  if (src->dtype != dst->dtype || src->width != dst->width || src->height != dst->height)
    return set_error(HIPACC_ERR_INVALID_ARG, "copy_region: windows differ in type or size");
  const int eb = dtype_bytes(src->dtype);
  hipStream_t s = static_cast<hipStream_t>(stream);
  for (int b = 0; b < batch_of(*src); ++b) {
    const char* sp = static_cast<const char*>(src->data) +
                     ((long long)b * src->plane_stride + (long long)src->offset_y * src->stride + src->offset_x) * eb;
    char* dp = static_cast<char*>(dst->data) +
               ((long long)b * dst->plane_stride + (long long)dst->offset_y * dst->stride + dst->offset_x) * eb;
    hipError_t e = hipMemcpy2DAsync(dp, (size_t)dst->stride * eb, sp, (size_t)src->stride * eb,
                                    (size_t)src->width * eb, src->height, hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) return hip_status(e, "copy_region");
  }
  return HIPACC_OK;
}

int hipacc_hip_fill(void* dst_dev, uint32_t value, size_t count, int32_t elem_bytes, void* stream) {
  if (!dst_dev) return set_error(HIPACC_ERR_INVALID_ARG, "fill: NULL pointer");
  return launch_fill(dst_dev, value, count, elem_bytes, static_cast<hipStream_t>(stream));
}

int hipacc_hip_fill_random(const hipacc_view* img, uint64_t seed, void* stream) {
  if (!img) return set_error(HIPACC_ERR_INVALID_ARG, "fill_random: NULL argument");
  int st = check_view(*img, "img");
  if (st) return st;
  return launch_random(img->data, img->dtype, img->stride, img->offset_x, img->offset_y, img->width,
                       img->height, batch_of(*img), img->plane_stride, seed,
                       static_cast<hipStream_t>(stream));
}

}  // extern "C"
