// reduce.hip -- global operators: reduced_data() and binned_data(), plus the
// small utility kernels (fill, synthetic pixels, window copy).
//
// Reduction: a fixed grid of blocks walks the window's rows (block b takes
// rows b, b+G, ...; threads stride along the row with 16-byte loads when the
// row is aligned, keeping the loads of 4 rows in flight together).  Each wave folds its 64 lanes with cross-lane shuffles,
// one LDS level folds the 4 waves, and block partials go to the workspace.
// A second one-block kernel folds the partials in a fixed order and writes
// the result, so the result does not depend on scheduling.  (A last-block
// ticket would need a device-scope release per block, which writes back
// the L2 of each XCD and costs more than the second launch.)
#include "hh_tile.hpp"

namespace hh {

template <typename A> __device__ __forceinline__ A red_identity(int mode);
template <> __device__ __forceinline__ float red_identity<float>(int mode) {
  switch (mode) {
    case HIPACC_REDUCE_MIN: return __builtin_huge_valf();
    case HIPACC_REDUCE_MAX: return -__builtin_huge_valf();
    case HIPACC_REDUCE_PROD: return 1.0f;
    default: return 0.0f;
  }
}
template <> __device__ __forceinline__ uint32_t red_identity<uint32_t>(int mode) {
  switch (mode) {
    case HIPACC_REDUCE_MIN: return 0xffffffffu;
    case HIPACC_REDUCE_MAX: return 0u;
    case HIPACC_REDUCE_PROD: return 1u;
    default: return 0u;
  }
}
template <> __device__ __forceinline__ int32_t red_identity<int32_t>(int mode) {
  switch (mode) {
    case HIPACC_REDUCE_MIN: return 0x7fffffff;
    case HIPACC_REDUCE_MAX: return (int32_t)0x80000000;
    case HIPACC_REDUCE_PROD: return 1;
    default: return 0;
  }
}

template <typename A> __device__ __forceinline__ A red_op(int mode, A a, A b) {
  switch (mode) {
    case HIPACC_REDUCE_MIN: return b < a ? b : a;
    case HIPACC_REDUCE_MAX: return b > a ? b : a;
    case HIPACC_REDUCE_PROD: return a * b;
    default: return a + b;
  }
}

template <typename A> __device__ __forceinline__ A wave_fold(int mode, A v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = red_op<A>(mode, v, __shfl_xor(v, off, 64));
  return v;
}

template <typename A> __device__ __forceinline__ A block_fold(int mode, A v, A* sm) {
  v = wave_fold<A>(mode, v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) sm[wave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    v = sm[0];
    for (int k = 1; k < 4; ++k) v = red_op<A>(mode, v, sm[k]);
  }
  return v;
}

// MODE is a template argument: a runtime mode would put a switch, with its
// branches, into the per-pixel loop.
template <typename T, typename A, int MODE>
__global__ __launch_bounds__(256) void reduce_kernel(const ReduceArgs a) {
  constexpr int VEC = 16 / (int)sizeof(T);
  constexpr int U = 4;   // rows whose loads a thread keeps in flight together
  __shared__ A sm[4];
  constexpr int mode = MODE;
  const int G = gridDim.x;
  A acc = red_identity<A>(mode);
  const T* base = static_cast<const T*>(a.in) + (long long)a.oy * a.stride + a.ox;
  if (a.vec_ok && a.stride == a.w) {
    // rows back to back: one grid-stride sweep, so that the whole grid reads
    // one contiguous stretch of memory at a time
    const long long nv = (long long)a.h * a.w / VEC;
    const long long st = (long long)G * 256;
    const uint4* v = reinterpret_cast<const uint4*>(base);
    long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    for (; i + (U - 1) * st < nv; i += U * st) {
      uint4 w[U];
#pragma unroll
      for (int u = 0; u < U; ++u) w[u] = v[i + u * st];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        T t[VEC];
        __builtin_memcpy(t, &w[u], 16);
#pragma unroll
        for (int k = 0; k < VEC; ++k) acc = red_op<A>(mode, acc, (A)t[k]);
      }
    }
    for (; i < nv; i += st) {
      const uint4 w = v[i];
      T t[VEC];
      __builtin_memcpy(t, &w, 16);
#pragma unroll
      for (int k = 0; k < VEC; ++k) acc = red_op<A>(mode, acc, (A)t[k]);
    }
    const long long n = (long long)a.h * a.w;
    for (long long x = nv * VEC + (long long)blockIdx.x * 256 + threadIdx.x; x < n; x += st)
      acc = red_op<A>(mode, acc, (A)base[x]);
  } else
  // block b folds rows b, b+G, ...; a thread loads the same vector of U of
  // its block's rows at once
  for (int y0 = blockIdx.x; y0 < a.h; y0 += U * G) {
    if (a.vec_ok) {
      const int nv = a.w / VEC;
      for (int i = threadIdx.x; i < nv; i += 256) {
        uint4 w[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int y = y0 + u * G;
          if (y < a.h) w[u] = reinterpret_cast<const uint4*>(base + (long long)y * a.stride)[i];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
          if (y0 + u * G >= a.h) break;
          T t[VEC];
          __builtin_memcpy(t, &w[u], 16);
#pragma unroll
          for (int k = 0; k < VEC; ++k) acc = red_op<A>(mode, acc, (A)t[k]);
        }
      }
      for (int u = 0; u < U; ++u) {
        const int y = y0 + u * G;
        if (y >= a.h) break;
        const T* row = base + (long long)y * a.stride;
        for (int x = nv * VEC + threadIdx.x; x < a.w; x += 256) acc = red_op<A>(mode, acc, (A)row[x]);
      }
    } else {
      for (int u = 0; u < U; ++u) {
        const int y = y0 + u * G;
        if (y >= a.h) break;
        const T* row = base + (long long)y * a.stride;
        for (int x = threadIdx.x; x < a.w; x += 256) acc = red_op<A>(mode, acc, (A)row[x]);
      }
    }
  }
  acc = block_fold<A>(mode, acc, sm);
  if (threadIdx.x == 0) {
    unsigned long long bits = 0;
    __builtin_memcpy(&bits, &acc, sizeof(A));
    a.partials[blockIdx.x] = bits;
  }
}

// Folds the n block partials in a fixed order (one block; the kernel
// boundary orders it after every partial is written).
template <typename A, int MODE>
__global__ __launch_bounds__(256) void reduce_final_kernel(const unsigned long long* partials,
                                                           int n, A* result) {
  __shared__ A sm[4];
  A v = red_identity<A>(MODE);
  for (int i = threadIdx.x; i < n; i += 256) {
    A p;
    __builtin_memcpy(&p, &partials[i], sizeof(A));
    v = red_op<A>(MODE, v, p);
  }
  v = block_fold<A>(MODE, v, sm);
  if (threadIdx.x == 0) *result = v;
}

template <typename T, typename A>
static int launch_reduce_t(const ReduceArgs& a, hipStream_t s) {
  const long long n = (long long)a.w * a.h;
  const int flat_blocks = (int)((n + 4095) / 4096 < kRedBlocks ? (n + 4095) / 4096 : kRedBlocks);
  const int blocks = (a.vec_ok && a.stride == a.w) ? (flat_blocks > 0 ? flat_blocks : 1)
                                                   : (a.h < kRedBlocks ? a.h : kRedBlocks);
#define HH_RED(M)                                                                         \
  hipLaunchKernelGGL((reduce_kernel<T, A, M>), dim3(blocks), dim3(256), 0, s, a);         \
  hipLaunchKernelGGL((reduce_final_kernel<A, M>), dim3(1), dim3(256), 0, s, a.partials, blocks, \
                     static_cast<A*>(a.result));
  switch (a.mode) {
    case HIPACC_REDUCE_MIN: HH_RED(HIPACC_REDUCE_MIN) break;
    case HIPACC_REDUCE_MAX: HH_RED(HIPACC_REDUCE_MAX) break;
    case HIPACC_REDUCE_PROD: HH_RED(HIPACC_REDUCE_PROD) break;
    default: HH_RED(HIPACC_REDUCE_SUM) break;
  }
#undef HH_RED
  return hip_status(hipGetLastError(), "reduce_kernel launch");
}

// acc type: float for F32 images; for integer images the requested
// accumulator (U32 or I32) -- min/max of integer images are exact in either.
int launch_reduce(const ReduceArgs& a, int dt, int acc_dt, hipStream_t s) {
  const bool ai = acc_dt == HIPACC_I32;
  switch (dt) {
    case HIPACC_F32:
      if (acc_dt != HIPACC_F32) break;
      return launch_reduce_t<float, float>(a, s);
    case HIPACC_U8: return ai ? launch_reduce_t<uint8_t, int32_t>(a, s) : launch_reduce_t<uint8_t, uint32_t>(a, s);
    case HIPACC_I8: return ai ? launch_reduce_t<int8_t, int32_t>(a, s) : launch_reduce_t<int8_t, uint32_t>(a, s);
    case HIPACC_U16: return ai ? launch_reduce_t<uint16_t, int32_t>(a, s) : launch_reduce_t<uint16_t, uint32_t>(a, s);
    case HIPACC_I16: return ai ? launch_reduce_t<int16_t, int32_t>(a, s) : launch_reduce_t<int16_t, uint32_t>(a, s);
    case HIPACC_U32: return ai ? launch_reduce_t<uint32_t, int32_t>(a, s) : launch_reduce_t<uint32_t, uint32_t>(a, s);
    case HIPACC_I32: return ai ? launch_reduce_t<int32_t, int32_t>(a, s) : launch_reduce_t<int32_t, uint32_t>(a, s);
  }
  return set_error(HIPACC_ERR_UNSUPPORTED, "reduce: image/accumulator type combination");
}

// ---------------------------------------------------------------------------
// Histogram.  Each block counts its share of the window into NC copies of an
// LDS histogram (bin b of copy c at word b * NC + c, a lane adds into copy
// lane % NC: with NC = 32 the 64 adds of a wave spread over all banks
// instead of piling onto the banks of the frequent bins), folds the copies
// and stores its bins as one row of the partials workspace; a second kernel,
// one block per bin, sums the rows.  Integer sums are exact in any order.  Merging with one global atomic per
// bin and block instead serialised the blocks' adds on each bin's address.
// ---------------------------------------------------------------------------
template <typename T> __device__ __forceinline__ int bin_of(T p, int nbins, float scale);
template <> __device__ __forceinline__ int bin_of<uint8_t>(uint8_t p, int nbins, float) {
  return (int)(((uint32_t)p * (uint32_t)nbins) >> 8);
}
template <> __device__ __forceinline__ int bin_of<uint16_t>(uint16_t p, int nbins, float) {
  return (int)(((uint32_t)p * (uint32_t)nbins) >> 16);
}
template <> __device__ __forceinline__ int bin_of<float>(float p, int nbins, float scale) {
  const float f = p * scale;
  if (!(f >= 0.0f) || f >= (float)nbins) return -1;
  return (int)(uint32_t)f;
}

constexpr int kHistThreads = 512;

template <typename T, int NC, bool B256>
__global__ __launch_bounds__(kHistThreads) void hist_partial_kernel(const HistArgs a) {
  constexpr int VEC = 16 / (int)sizeof(T);
  constexpr int NT = kHistThreads;
  extern __shared__ uint32_t hist[];
  const int nb = B256 ? 256 : a.nbins;
  for (int i = threadIdx.x; i < nb * NC; i += NT) hist[i] = 0u;
  __syncthreads();
  uint32_t* mine = hist + (threadIdx.x & (NC - 1));
  auto count1 = [&](T p) {
    if constexpr (B256) {
      atomicAdd(mine + (int)p * NC, 1u);
    } else {
      const int b = bin_of<T>(p, a.nbins, a.scale);
      if (b >= 0 && b < a.nbins) atomicAdd(mine + b * NC, 1u);
    }
  };
  auto count = [&](const uint4& w) {
    T t[VEC];
    __builtin_memcpy(t, &w, 16);
#pragma unroll
    for (int k = 0; k < VEC; ++k) count1(t[k]);
  };
  if (a.vec_ok && a.stride == a.w) {
    // Rows back to back: block b takes the contiguous vectors
    // [b*nv/G, (b+1)*nv/G).  Each thread keeps the loads of its next U
    // vectors in flight while it counts the current U.
    constexpr int U = 4;
    const T* base = static_cast<const T*>(a.in) + (long long)a.oy * a.stride + a.ox;
    const long long n = (long long)a.h * a.w;
    const long long nv = n / VEC;
    const long long b0 = nv * blockIdx.x / gridDim.x, b1 = nv * (blockIdx.x + 1) / gridDim.x;
    const uint4* v = reinterpret_cast<const uint4*>(base);
    long long i = b0 + threadIdx.x;
    if (i + (U - 1) * NT < b1) {
      uint4 cur[U];
#pragma unroll
      for (int u = 0; u < U; ++u) cur[u] = v[i + u * NT];
      for (i += U * NT; i + (U - 1) * NT < b1; i += U * NT) {
        uint4 nxt[U];
#pragma unroll
        for (int u = 0; u < U; ++u) nxt[u] = v[i + u * NT];
#pragma unroll
        for (int u = 0; u < U; ++u) count(cur[u]);
#pragma unroll
        for (int u = 0; u < U; ++u) cur[u] = nxt[u];
      }
#pragma unroll
      for (int u = 0; u < U; ++u) count(cur[u]);
    }
    for (; i < b1; i += NT) count(v[i]);
    if (blockIdx.x == gridDim.x - 1)
      for (long long x = nv * VEC + threadIdx.x; x < n; x += NT) count1(base[x]);
  } else {
    for (int y = blockIdx.x; y < a.h; y += gridDim.x) {
      const T* row = static_cast<const T*>(a.in) + (long long)(a.oy + y) * a.stride + a.ox;
      int start = 0;
      if (a.vec_ok) {
        const int nv = a.w / VEC;
        for (int i = threadIdx.x; i < nv; i += NT) count(reinterpret_cast<const uint4*>(row)[i]);
        start = nv * VEC;
      }
      for (int x = start + threadIdx.x; x < a.w; x += NT) count1(row[x]);
    }
  }
  __syncthreads();
  // fold the copies (copy c of bin b read as (b + c) % NC: a wave's lanes,
  // consecutive bins, read consecutive banks) into this block's row
  uint32_t* part = a.partials + (long long)blockIdx.x * nb;
  for (int b = threadIdx.x; b < nb; b += NT) {
    uint32_t s = 0;
#pragma unroll
    for (int c = 0; c < NC; ++c) s += hist[b * NC + ((b + c) & (NC - 1))];
    part[b] = s;
  }
}

__global__ __launch_bounds__(256) void hist_merge_kernel(const uint32_t* partials, int nparts,
                                                         int nbins, uint32_t* bins) {
  __shared__ uint32_t red[4];
  const int b = blockIdx.x;
  uint32_t s = 0;
  for (int p = threadIdx.x; p < nparts; p += 256) s += partials[(long long)p * nbins + b];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) bins[b] = (red[0] + red[1]) + (red[2] + red[3]);
}

template <typename T, int NC, bool B256>
static int launch_hist_t(const HistArgs& a, int blocks, hipStream_t s) {
  const size_t smem = (size_t)(B256 ? 256 : a.nbins) * NC * 4;
  hipLaunchKernelGGL((hist_partial_kernel<T, NC, B256>), dim3(blocks), dim3(kHistThreads), smem, s,
                     a);
  return hip_status(hipGetLastError(), "hist_partial_kernel launch");
}

int launch_histogram(const HistArgs& a, int dt, hipStream_t s) {
  const int eb = dt == HIPACC_U8 ? 1 : dt == HIPACC_U16 ? 2 : 4;
  const long long n = (long long)a.w * a.h;
  // Contiguous windows: about 64 KiB per block (4096^2 u8: 256 blocks,
  // 9.9 us for both launches; 16, 32 and 128 KiB: 11.2-11.4 us), at most 512
  // blocks; other windows: one block per row.  blocks * nbins entries must
  // fit the partials workspace.
  const bool flat = a.vec_ok && a.stride == a.w;
  const long long chunk = 64 * 1024;
  const long long want = flat ? (n * eb + chunk - 1) / chunk : a.h;
  long long blocks = want < 512 ? want : 512;
  const long long cap = kHistPartialEntries / a.nbins;
  if (blocks > cap) blocks = cap;
  if (blocks < 1) blocks = 1;
  int st;
  switch (dt) {
    case HIPACC_U8:
      st = a.nbins == 256 ? launch_hist_t<uint8_t, 32, true>(a, (int)blocks, s)
           : a.nbins <= 256 ? launch_hist_t<uint8_t, 32, false>(a, (int)blocks, s)
           : a.nbins <= 1024 ? launch_hist_t<uint8_t, 8, false>(a, (int)blocks, s)
                             : launch_hist_t<uint8_t, 1, false>(a, (int)blocks, s);
      break;
    case HIPACC_U16:
      st = a.nbins <= 256 ? launch_hist_t<uint16_t, 32, false>(a, (int)blocks, s)
           : a.nbins <= 1024 ? launch_hist_t<uint16_t, 8, false>(a, (int)blocks, s)
                             : launch_hist_t<uint16_t, 1, false>(a, (int)blocks, s);
      break;
    case HIPACC_F32:
      st = a.nbins <= 256 ? launch_hist_t<float, 32, false>(a, (int)blocks, s)
           : a.nbins <= 1024 ? launch_hist_t<float, 8, false>(a, (int)blocks, s)
                             : launch_hist_t<float, 1, false>(a, (int)blocks, s);
      break;
    default:
      return set_error(HIPACC_ERR_UNSUPPORTED, "histogram: pixel type must be U8, U16 or F32");
  }
  if (st) return st;
  hipLaunchKernelGGL(hist_merge_kernel, dim3(a.nbins), dim3(256), 0, s, a.partials, (int)blocks,
                     a.nbins, a.bins);
  return hip_status(hipGetLastError(), "hist_merge_kernel launch");
}

// ---------------------------------------------------------------------------
// Utilities
// ---------------------------------------------------------------------------
template <typename W>
__global__ void fill_kernel(W* p, W v, size_t n) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) p[i] = v;
}

int launch_fill(void* p, uint32_t value, size_t n, int eb, hipStream_t s) {
  if (n == 0) return HIPACC_OK;
  size_t blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  switch (eb) {
    case 1: hipLaunchKernelGGL(fill_kernel<uint8_t>, dim3(blocks), dim3(256), 0, s, (uint8_t*)p, (uint8_t)value, n); break;
    case 2: hipLaunchKernelGGL(fill_kernel<uint16_t>, dim3(blocks), dim3(256), 0, s, (uint16_t*)p, (uint16_t)value, n); break;
    case 4: hipLaunchKernelGGL(fill_kernel<uint32_t>, dim3(blocks), dim3(256), 0, s, (uint32_t*)p, value, n); break;
    default: return set_error(HIPACC_ERR_INVALID_ARG, "fill: elem_bytes must be 1, 2 or 4");
  }
  return hip_status(hipGetLastError(), "fill_kernel launch");
}

// SplitMix64 (Steele, Lea, Flood 2014): the k-th output of a generator seeded
// with s is mix(s + k * 0x9e3779b97f4a7c15), k >= 1.
HH_HD uint64_t splitmix64_at(uint64_t seed, uint64_t k) {
  uint64_t z = seed + k * 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

template <typename T> __device__ __forceinline__ T rand_pixel(uint64_t x);
template <> __device__ __forceinline__ uint8_t rand_pixel<uint8_t>(uint64_t x) { return (uint8_t)(x >> 56); }
template <> __device__ __forceinline__ int8_t rand_pixel<int8_t>(uint64_t x) { return (int8_t)(x >> 56); }
template <> __device__ __forceinline__ uint16_t rand_pixel<uint16_t>(uint64_t x) { return (uint16_t)(x >> 48); }
template <> __device__ __forceinline__ int16_t rand_pixel<int16_t>(uint64_t x) { return (int16_t)(x >> 48); }
template <> __device__ __forceinline__ uint32_t rand_pixel<uint32_t>(uint64_t x) { return (uint32_t)(x >> 32); }
template <> __device__ __forceinline__ int32_t rand_pixel<int32_t>(uint64_t x) { return (int32_t)(x >> 32); }
template <> __device__ __forceinline__ float rand_pixel<float>(uint64_t x) {
  return (float)(x >> 40) * (1.0f / 16777216.0f);
}

template <typename T>
__global__ void random_kernel(T* p, int stride, int ox, int oy, int w, int h, long long plane,
                              uint64_t seed) {
  const int x = blockIdx.x * 256 + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= w) return;
  const uint64_t k = (uint64_t)blockIdx.z * (uint64_t)w * (uint64_t)h + (uint64_t)y * w + x + 1;
  p[(long long)blockIdx.z * plane + (long long)(oy + y) * stride + ox + x] = rand_pixel<T>(splitmix64_at(seed, k));
}

int launch_random(void* p, int dt, int stride, int ox, int oy, int w, int h, int batch, long long plane,
                  uint64_t seed, hipStream_t s) {
  dim3 grid(ceil_div(w, 256), h, batch);
#define HH_RND(T) hipLaunchKernelGGL(random_kernel<T>, grid, dim3(256), 0, s, (T*)p, stride, ox, oy, w, h, plane, seed)
  switch (dt) {
    case HIPACC_U8: HH_RND(uint8_t); break;
    case HIPACC_I8: HH_RND(int8_t); break;
    case HIPACC_U16: HH_RND(uint16_t); break;
    case HIPACC_I16: HH_RND(int16_t); break;
    case HIPACC_U32: HH_RND(uint32_t); break;
    case HIPACC_I32: HH_RND(int32_t); break;
    case HIPACC_F32: HH_RND(float); break;
    default: return set_error(HIPACC_ERR_INVALID_ARG, "fill_random: bad dtype");
  }
#undef HH_RND
  return hip_status(hipGetLastError(), "random_kernel launch");
}

}  // namespace hh
