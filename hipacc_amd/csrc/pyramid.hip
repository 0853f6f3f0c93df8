// pyramid.hip -- resampling point operators (Interpolate::NN / LF) and the
// batched Gaussian / Laplacian pyramid: one launch per level over a work list
// of tiles of every image (one block per tile), the launches captured once
// into a hipGraph.
#include <hip/hip_runtime.h>

#include <cstring>
#include <type_traits>
#include <vector>

#include "hh_pyramid.hpp"

namespace hh {

// ---------------------------------------------------------------------------
// Resampling point operator
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ float rs_read(const ResampleArgs& a, int x, int y) {
  if (a.boundary == HIPACC_BOUNDARY_CONSTANT) {
    if (x < a.lo_x || x >= a.hi_x || y < a.lo_y || y >= a.hi_y) return a.cval;
  } else {
    x = bh_map(a.boundary, x, a.lo_x, a.hi_x, a.in_imgw);
    y = bh_map(a.boundary, y, a.lo_y, a.hi_y, a.in_imgh);
  }
  return (float)static_cast<const T*>(a.in)[(long long)y * a.in_stride + x];
}

// Interpolation::interpolate (dsl/image.hpp:390-532) at iteration-space
// pixel (x, y) read at (x + xf, y + yf).
template <typename T, int MODE>
__device__ __forceinline__ float rs_value(const ResampleArgs& a, int x, int y) {
  const float mx = interp_map(a.sx, x + a.xf, a.lo_x);
  const float my = interp_map(a.sy, y + a.yf, a.lo_y);
  return interp_sample<MODE>([&](int px, int py) { return rs_read<T>(a, px, py); }, mx, my);
}

template <typename T, int MODE>
__global__ __launch_bounds__(256) void resample_kernel(const ResampleArgs a) {
  const int x = blockIdx.x * 64 + (threadIdx.x & 63);
  const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (x >= a.w || y >= a.h) return;
  float v = rs_value<T, MODE>(a, x, y);
  if (a.op != HIPACC_RESAMPLE_COPY) {
    const float b = a.in2[(long long)(y + a.in2_oy) * a.in2_stride + x + a.in2_ox];
    v = a.op == HIPACC_RESAMPLE_SUB ? b - v : v + b;
  }
  const long long o = (long long)(y + a.out_oy) * a.out_stride + x + a.out_ox;
  if constexpr (std::is_same<T, float>::value) static_cast<float*>(a.out)[o] = v;
  else store_f(a.out, a.out_dtype, o, v, HIPACC_CVT_TRUNC, 0.0f);
}

template <typename T>
static int launch_resample_t(const ResampleArgs& a, dim3 grid, hipStream_t s) {
  switch (a.interp) {
    case HIPACC_INTERP_NN:
      hipLaunchKernelGGL((resample_kernel<T, HIPACC_INTERP_NN>), grid, dim3(256), 0, s, a); break;
    case HIPACC_INTERP_LF:
      hipLaunchKernelGGL((resample_kernel<T, HIPACC_INTERP_LF>), grid, dim3(256), 0, s, a); break;
    case HIPACC_INTERP_B5:
      hipLaunchKernelGGL((resample_kernel<T, HIPACC_INTERP_B5>), grid, dim3(256), 0, s, a); break;
    case HIPACC_INTERP_CF:
      hipLaunchKernelGGL((resample_kernel<T, HIPACC_INTERP_CF>), grid, dim3(256), 0, s, a); break;
    case HIPACC_INTERP_L3:
      hipLaunchKernelGGL((resample_kernel<T, HIPACC_INTERP_L3>), grid, dim3(256), 0, s, a); break;
    default: return set_error(HIPACC_ERR_INVALID_ARG, "resample: interp %d", a.interp);
  }
  return hip_status(hipGetLastError(), "resample_kernel launch");
}

int launch_resample(const ResampleArgs& a, hipStream_t s) {
  dim3 grid(ceil_div(a.w, 64), ceil_div(a.h, 4));
  return a.in_dtype == HIPACC_U8 ? launch_resample_t<uint8_t>(a, grid, s)
                                 : launch_resample_t<float>(a, grid, s);
}

// ---------------------------------------------------------------------------
// Pyramid level kernel
// ---------------------------------------------------------------------------
// G[l] read with the convolution's boundary mode over the whole image.
__device__ __forceinline__ float pyr_read(const float* g, int stride, int w, int h, int mode,
                                          int x, int y) {
  if (x < 0 || x >= w || y < 0 || y >= h) {
    if (mode == HIPACC_BOUNDARY_CONSTANT) return 0.0f;
    x = bh_map(mode, x, 0, w, w);
    y = bh_map(mode, y, 0, h, h);
  }
  return g[(long long)y * stride + x];
}

// For each tile (TW x TH pixels of level l): the G[l] region the tile reads
// is staged in LDS once (with the convolution's boundary mode); from it the
// G[l+1] values the tile's LF reads and the ones it owns (columns x0/2..,
// rows y0/2..) are computed -- the 5x5 convolution evaluated only at the NN
// sample points -- the owned ones are stored, then L[l] = G[l] - LF(G[l+1]).
// SNUM / 4 bounds the NN step s = W / W1 of every image in the work list:
// 12 (s <= 3, any image) or 9 (s <= 2.25: images at least 9 pixels wide and
// high at this level), which sizes the staged region.
template <int TW, int TH, int SNUM, bool FMA>
__global__ __launch_bounds__(256) void pyramid_level_kernel(const PyrLevelArgs a) {
  using Cap = PyrCap<TW, TH, SNUM>;
  constexpr int UW = Cap::UW, UH = Cap::UH, RW = Cap::RW, RH = Cap::RH;
  __shared__ float g1s[UH * UW];
  __shared__ float gls[RH * RW];
  const int l = a.level;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;

  for (int t = blockIdx.x; t < a.ntiles; t += gridDim.x) {
    const PyrTile tile = a.tiles[t];
    const PyrImage& im = a.imgs[tile.img];
    const int W = im.w[l], H = im.h[l], W1 = im.w[l + 1], H1 = im.h[l + 1];
    const float* G = im.g[l];
    float* G1 = im.g[l + 1];
    float* L = im.lap[l];
    const int gs = im.gs[l], g1st = im.gs[l + 1], ls = im.ls[l];
    const int x0 = tile.x0, y0 = tile.y0;
    const float su = (float)W1 / (float)W, sv = (float)H1 / (float)H;
    const float sd = (float)W / (float)W1, td = (float)H / (float)H1;
    const PyrRegion rg = pyr_tile_region(W, H, W1, H1, x0, y0, TW, TH);
    const int ux0 = rg.ux0, uy0 = rg.uy0, uw = rg.uw, uh = rg.uh;
    const int ox0 = rg.ox0, ox1 = rg.ox1, oy0 = rg.oy0, oy1 = rg.oy1;
    const int rx0 = rg.rx0, ry0 = rg.ry0, rw = rg.rw, rh = rg.rh;
    const int rx1 = rx0 + rw - 1, ry1 = ry0 + rh - 1;
    // The plan verified on the host (pyr_plan_tiles) that every tile of its
    // work lists fits; this guard only keeps LDS accesses in bounds.
    if (!Cap::fits(rg)) continue;

    __syncthreads();  // previous tile's LDS reads are done
    // wave w stages rows w, w+4, ...; lane c the columns c, c+64, ...  Inside
    // the image every load of the region is issued before the first LDS
    // write, so the whole region is in flight at once.
    if (rx0 >= 0 && ry0 >= 0 && rx1 < W && ry1 < H) {
      constexpr int NR = (RH + 3) / 4, NC = (RW + 63) / 64;
      float v[NR][NC];
#pragma unroll
      for (int k = 0; k < NR; ++k) {
        const int r = wave + 4 * k;
        const float* src = G + (long long)(ry0 + min(r, rh - 1)) * gs + rx0;
#pragma unroll
        for (int c = 0; c < NC; ++c) v[k][c] = src[min(lane + 64 * c, rw - 1)];
      }
#pragma unroll
      for (int k = 0; k < NR; ++k) {
        const int r = wave + 4 * k;
#pragma unroll
        for (int c = 0; c < NC; ++c)
          if (r < rh && lane + 64 * c < rw) gls[r * RW + lane + 64 * c] = v[k][c];
      }
    } else {
      for (int r = wave; r < rh; r += 4)
        for (int c = lane; c < rw; c += 64)
          gls[r * RW + c] = pyr_read(G, gs, W, H, a.boundary, rx0 + c, ry0 + r);
    }
    __syncthreads();
    // i / uw as a multiply by the rounded-up reciprocal ceil(2^32 / uw):
    // exact for i < 2^32 / uw (i < UW * UH here); uw == 1 has no 32-bit
    // reciprocal and takes i itself
    const unsigned recip = uw > 1 ? 0xffffffffu / (unsigned)uw + 1u : 0u;
    for (int i = threadIdx.x; i < uw * uh; i += 256) {
      const int yy = uw > 1 ? (int)__umulhi((unsigned)i, recip) : i, xx = i - yy * uw;
      const int xu = ux0 + xx, yu = uy0 + yy;
      const int nx = (int)interp_map(sd, xu) - 2 - rx0, ny = (int)interp_map(td, yu) - 2 - ry0;
      const float* src = gls + ny * RW + nx;
      float acc = 0.0f;
#pragma unroll
      for (int ty = 0; ty < 5; ++ty)
#pragma unroll
        for (int tx = 0; tx < 5; ++tx) {
          const float v = src[ty * RW + tx];
          const float wv = a.w[ty * 5 + tx];
          if (ty == 0 && tx == 0)
            acc = mul_first(wv, v);
          else
            acc = madd<FMA>(wv, v, acc);
        }
      g1s[yy * UW + xx] = acc;
      if (xu >= ox0 && xu <= ox1 && yu >= oy0 && yu <= oy1)
        G1[(long long)yu * g1st + xu] = acc;
    }
    __syncthreads();

    // a lane keeps the columns lane, lane+64, ... (their LF column weights
    // are computed once) and walks the rows wave, wave+4, ... (the row's LF
    // weights computed once for all its columns)
    constexpr int NCX = TW / 64;
    int xa[NCX], xb[NCX];
    float fx[NCX];
#pragma unroll
    for (int cx = 0; cx < NCX; ++cx) {
      const int x = x0 + lane + 64 * cx;
      int xi;
      interp_lf(interp_map(su, x), xi, fx[cx]);
      xa[cx] = xi - ux0;
      xb[cx] = min(xi + 1, W1 - 1) - ux0;
    }
    const int rows = min(TH, H - y0);
    for (int r = wave; r < rows; r += 4) {
      const int y = y0 + r;
      int yi;
      float fy;
      interp_lf(interp_map(sv, y), yi, fy);
      const int ya = (yi - uy0) * UW, yb = (min(yi + 1, H1 - 1) - uy0) * UW;
      const int grow = (y - ry0) * RW - rx0;
      float* lrow = L + (long long)y * ls;
#pragma unroll
      for (int cx = 0; cx < NCX; ++cx) {
        const int x = x0 + lane + 64 * cx;
        if (x < W) {
          const float lf = interp_bilinear(fx[cx], fy, g1s[ya + xa[cx]], g1s[ya + xb[cx]],
                                           g1s[yb + xa[cx]], g1s[yb + xb[cx]]);
          // L[l] is not read again here: non-temporal stores
          __builtin_nontemporal_store(gls[grow + x] - lf, lrow + x);
        }
      }
    }
  }
}

}  // namespace hh

// ---------------------------------------------------------------------------
// Plans
// ---------------------------------------------------------------------------
struct hipacc_pyramid_plan {
  int levels = 0, nimg = 0, tile_rows = 16, fma = 1, boundary = HIPACC_BOUNDARY_CLAMP;
  float w[25];
  hh::PyrImage* d_imgs = nullptr;
  hh::PyrTile* d_tiles = nullptr;
  std::vector<int> off, cnt;     // narrow tiles (64 x tile_rows) per level
  std::vector<int> offw, cntw;   // wide tiles (128 x 32) per level
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace hh {

constexpr int kPyrWideMin = 512;    // wide tiles of a level worth a wide launch

template <int TW, int TH, int SNUM, bool FMA>
static int launch_level_t(const PyrLevelArgs& a, hipStream_t s) {
  // one block per tile: the hardware keeps several tiles per CU in flight,
  // hiding the staging latency of one behind the arithmetic of another
  hipLaunchKernelGGL((pyramid_level_kernel<TW, TH, SNUM, FMA>), dim3(a.ntiles), dim3(256), 0, s, a);
  return hip_status(hipGetLastError(), "pyramid_level_kernel launch");
}

template <bool FMA>
static int launch_level(const PyrLevelArgs& a, bool wide, int tile_rows, hipStream_t s) {
  if (wide) return launch_level_t<128, 32, 9, FMA>(a, s);
  return tile_rows == 8    ? launch_level_t<64, 8, 12, FMA>(a, s)
         : tile_rows == 32 ? launch_level_t<64, 32, 12, FMA>(a, s)
                           : launch_level_t<64, 16, 12, FMA>(a, s);
}

static int plan_launches(hipacc_pyramid_plan* p, hipStream_t s) {
  for (int l = 0; l + 1 < p->levels; ++l) {
    for (int wide = 1; wide >= 0; --wide) {
      const int n = wide ? p->cntw[l] : p->cnt[l];
      if (n == 0) continue;
      PyrLevelArgs a;
      std::memset(&a, 0, sizeof(a));
      a.imgs = p->d_imgs;
      a.tiles = p->d_tiles + (wide ? p->offw[l] : p->off[l]);
      a.ntiles = n;
      a.level = l;
      a.boundary = p->boundary;
      std::memcpy(a.w, p->w, sizeof(a.w));
      const int st = p->fma ? launch_level<true>(a, wide != 0, p->tile_rows, s)
                            : launch_level<false>(a, wide != 0, p->tile_rows, s);
      if (st) return st;
    }
  }
  return HIPACC_OK;
}

// Work lists of every level: images at least 9 x 9 at a level (NN step
// <= 2.25) may take wide 128 x 32 tiles; a level takes them when there are
// at least kPyrWideMin (they fill the device several times over), and the
// smaller images of that level, or every image of a smaller level, take
// narrow 64 x tile_rows tiles.  Every tile's staged regions are computed with
// the kernel's own arithmetic and checked against the LDS capacity of the
// kernel that will run it; a tile that does not fit is an error, never a
// skipped tile.
static int pyr_plan_tiles(const int* w0, const int* h0, int n, int levels, int tile_rows,
                          std::vector<std::vector<PyrTile>>& lists,
                          std::vector<std::vector<PyrTile>>& wlists) {
  lists.assign(levels, {});
  wlists.assign(levels, {});
  std::vector<std::vector<PyrTile>> nlists(levels);
  std::vector<int> W(n), H(n);
  for (int i = 0; i < n; ++i) {
    int w = w0[i], h = h0[i];
    for (int l = 0; l < levels; ++l) {
      if (w < 1 || h < 1)
        return set_error(HIPACC_ERR_INVALID_ARG, "pyramid: image %d too small for %d levels", i,
                         levels);
      if (l + 1 < levels) {
        if (w >= 9 && h >= 9)
          for (int y0 = 0; y0 < h; y0 += 32)
            for (int x0 = 0; x0 < w; x0 += 128) wlists[l].push_back(PyrTile{i, x0, y0});
        for (int y0 = 0; y0 < h; y0 += tile_rows)
          for (int x0 = 0; x0 < w; x0 += 64) nlists[l].push_back(PyrTile{i, x0, y0});
      }
      w /= 2;
      h /= 2;
    }
  }
  auto lw = [&](int i, int l) { return w0[i] >> l; };
  auto lh = [&](int i, int l) { return h0[i] >> l; };
  for (int l = 0; l + 1 < levels; ++l) {
    if ((int)wlists[l].size() >= kPyrWideMin) {
      for (const PyrTile& t : nlists[l])
        if (lw(t.img, l) < 9 || lh(t.img, l) < 9) lists[l].push_back(t);
    } else {
      wlists[l].clear();
      lists[l] = nlists[l];
    }
    for (int wide = 0; wide < 2; ++wide)
      for (const PyrTile& t : wide ? wlists[l] : lists[l]) {
        const int Wl = lw(t.img, l), Hl = lh(t.img, l);
        const int TW = wide ? 128 : 64, TH = wide ? 32 : tile_rows;
        const PyrRegion r = pyr_tile_region(Wl, Hl, Wl / 2, Hl / 2, t.x0, t.y0, TW, TH);
        const bool ok = wide ? PyrCap<128, 32, 9>::fits(r)
                        : tile_rows == 8  ? PyrCap<64, 8, 12>::fits(r)
                        : tile_rows == 32 ? PyrCap<64, 32, 12>::fits(r)
                                          : PyrCap<64, 16, 12>::fits(r);
        if (!ok)
          return set_error(HIPACC_ERR_INVALID_ARG,
                           "pyramid: image %d level %d tile (%d, %d) stages %dx%d / %dx%d "
                           "pixels, more than its kernel holds",
                           t.img, l, t.x0, t.y0, r.uw, r.uh, r.rw, r.rh);
      }
  }
  return HIPACC_OK;
}

static void plan_free(hipacc_pyramid_plan* p) {
  if (!p) return;
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  if (p->d_imgs) (void)hipFree(p->d_imgs);
  if (p->d_tiles) (void)hipFree(p->d_tiles);
  delete p;
}

}  // namespace hh

using namespace hh;

extern "C" {

int hipacc_hip_pyramid_plan_create(const hipacc_pyramid_desc* d, void* stream,
                                   hipacc_pyramid_plan** out) {
  if (!d || !out) return set_error(HIPACC_ERR_INVALID_ARG, "pyramid: NULL argument");
  *out = nullptr;
  if (d->num_images < 1 || !d->images)
    return set_error(HIPACC_ERR_INVALID_ARG, "pyramid: no images");
  if (d->levels < 2 || d->levels > HIPACC_MAX_LEVELS)
    return set_error(HIPACC_ERR_INVALID_ARG, "pyramid: levels %d (2..%d)", d->levels,
                     HIPACC_MAX_LEVELS);
  if (d->boundary < HIPACC_BOUNDARY_UNDEFINED || d->boundary > HIPACC_BOUNDARY_CONSTANT)
    return set_error(HIPACC_ERR_INVALID_ARG, "pyramid: boundary %d", d->boundary);
  if (d->tile_rows != 0 && d->tile_rows != 8 && d->tile_rows != 16 && d->tile_rows != 32)
    return set_error(HIPACC_ERR_INVALID_ARG, "pyramid: tile_rows %d (0, 8, 16, 32)", d->tile_rows);
  hipacc_pyramid_plan* p = new hipacc_pyramid_plan;
  p->levels = d->levels;
  p->nimg = d->num_images;
  p->tile_rows = d->tile_rows ? d->tile_rows : 16;
  p->fma = d->fp_contract != 0;
  p->boundary = d->boundary;
  static const float b5[5] = {1.0f, 4.0f, 6.0f, 4.0f, 1.0f};
  for (int t = 0; t < 25; ++t) p->w[t] = d->mask ? d->mask[t] : b5[t / 5] * b5[t % 5] / 256.0f;

  std::vector<PyrImage> imgs(d->num_images);
  p->off.assign(d->levels, 0);
  p->cnt.assign(d->levels, 0);
  p->offw.assign(d->levels, 0);
  p->cntw.assign(d->levels, 0);
  std::vector<int> w0(d->num_images), h0(d->num_images);
  for (int i = 0; i < d->num_images; ++i) {
    w0[i] = d->images[i].width;
    h0[i] = d->images[i].height;
  }
  std::vector<std::vector<PyrTile>> lists, wlists;
  int st0 = pyr_plan_tiles(w0.data(), h0.data(), d->num_images, d->levels, p->tile_rows, lists,
                           wlists);
  if (st0) {
    plan_free(p);
    return st0;
  }
  for (int i = 0; i < d->num_images; ++i) {
    const hipacc_pyramid_image& s = d->images[i];
    PyrImage& m = imgs[i];
    std::memset(&m, 0, sizeof(m));
    int w = s.width, h = s.height;
    for (int l = 0; l < d->levels; ++l) {
      if (!s.gauss[l] || s.gauss_stride[l] < w || (l + 1 < d->levels && (!s.lap[l] || s.lap_stride[l] < w))) {
        plan_free(p);
        return set_error(HIPACC_ERR_INVALID_ARG, "pyramid: image %d level %d buffers", i, l);
      }
      m.w[l] = w;
      m.h[l] = h;
      m.g[l] = s.gauss[l];
      m.gs[l] = s.gauss_stride[l];
      if (l + 1 < d->levels) {
        m.lap[l] = s.lap[l];
        m.ls[l] = s.lap_stride[l];
      }
      w /= 2;
      h /= 2;
    }
  }
  std::vector<PyrTile> all;
  for (int l = 0; l < d->levels; ++l) {
    p->off[l] = (int)all.size();
    p->cnt[l] = (int)lists[l].size();
    all.insert(all.end(), lists[l].begin(), lists[l].end());
    p->offw[l] = (int)all.size();
    p->cntw[l] = (int)wlists[l].size();
    all.insert(all.end(), wlists[l].begin(), wlists[l].end());
  }
  hipError_t e = hipMalloc(&p->d_imgs, sizeof(PyrImage) * imgs.size());
  if (e == hipSuccess) e = hipMalloc(&p->d_tiles, sizeof(PyrTile) * (all.empty() ? 1 : all.size()));
  hipStream_t us = static_cast<hipStream_t>(stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(p->d_imgs, imgs.data(), sizeof(PyrImage) * imgs.size(),
                       hipMemcpyHostToDevice, us);
  if (e == hipSuccess && !all.empty())
    e = hipMemcpyAsync(p->d_tiles, all.data(), sizeof(PyrTile) * all.size(),
                       hipMemcpyHostToDevice, us);
  if (e == hipSuccess) e = hipStreamSynchronize(us);
  if (e != hipSuccess) {
    plan_free(p);
    return hip_status(e, "pyramid: work list upload");
  }
  // Capture the level launches once on a private stream.
  hipStream_t cs = nullptr;
  e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
  if (e != hipSuccess) {
    plan_free(p);
    return hip_status(e, "hipStreamCreateWithFlags");
  }
  int st = HIPACC_OK;
  e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    st = plan_launches(p, cs);
    hipError_t e2 = hipStreamEndCapture(cs, &p->graph);
    if (st == HIPACC_OK && e2 != hipSuccess) st = hip_status(e2, "hipStreamEndCapture");
  } else {
    st = hip_status(e, "hipStreamBeginCapture");
  }
  if (st == HIPACC_OK) {
    e = hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) st = hip_status(e, "hipGraphInstantiate");
  }
  (void)hipStreamDestroy(cs);
  if (st) {
    plan_free(p);
    return st;
  }
  *out = p;
  return HIPACC_OK;
}

int hipacc_hip_pyramid_plan_tiles(const int32_t* widths, const int32_t* heights,
                                  int32_t num_images, int32_t levels, int32_t tile_rows,
                                  int32_t* narrow_tiles, int32_t* wide_tiles) {
  if (!widths || !heights || num_images < 1)
    return set_error(HIPACC_ERR_INVALID_ARG, "pyramid_plan_tiles: no images");
  if (levels < 2 || levels > HIPACC_MAX_LEVELS)
    return set_error(HIPACC_ERR_INVALID_ARG, "pyramid_plan_tiles: levels %d", levels);
  if (tile_rows == 0) tile_rows = 16;
  if (tile_rows != 8 && tile_rows != 16 && tile_rows != 32)
    return set_error(HIPACC_ERR_INVALID_ARG, "pyramid_plan_tiles: tile_rows %d", tile_rows);
  std::vector<std::vector<PyrTile>> lists, wlists;
  const int st = pyr_plan_tiles(widths, heights, num_images, levels, tile_rows, lists, wlists);
  if (st) return st;
  for (int l = 0; l + 1 < levels; ++l) {
    if (narrow_tiles) narrow_tiles[l] = (int32_t)lists[l].size();
    if (wide_tiles) wide_tiles[l] = (int32_t)wlists[l].size();
  }
  return HIPACC_OK;
}

int hipacc_hip_pyramid_run(hipacc_pyramid_plan* p, void* stream) {
  if (!p || !p->exec) return set_error(HIPACC_ERR_INVALID_ARG, "pyramid_run: no plan");
  return hip_status(hipGraphLaunch(p->exec, static_cast<hipStream_t>(stream)), "hipGraphLaunch");
}

int hipacc_hip_pyramid_run_direct(hipacc_pyramid_plan* p, void* stream) {
  if (!p) return set_error(HIPACC_ERR_INVALID_ARG, "pyramid_run_direct: no plan");
  return plan_launches(p, static_cast<hipStream_t>(stream));
}

int hipacc_hip_pyramid_plan_destroy(hipacc_pyramid_plan* p) {
  plan_free(p);
  return HIPACC_OK;
}

}  // extern "C"
