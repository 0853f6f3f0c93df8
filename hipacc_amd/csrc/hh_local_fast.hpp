// hh_local_fast.hpp -- the templated weighted-SUM local operator
// (Kernel::convolve with Reduce::SUM, Kernel::reduce(Domain, SUM)).
//
// Each lane owns VEC adjacent output columns and PPT output rows of its wave's
// quarter of the tile.  Input rows are streamed top to bottom from LDS; every
// row is widened once and added into the accumulators of all output rows
// whose window contains it, so each output pixel receives its taps in
// row-major order (dy outer, dx inner) -- the order of
// lib/AST/Convolution.cpp:397-435 and of the DSL's ElementIterator.
#pragma once

#include "hh_tile.hpp"

namespace hh {

template <typename TI, typename TO, typename ACC, int MW, int MH, int PPT, bool FMA, bool WDEV>
__global__ __launch_bounds__(256) void local_sum_kernel(const LocalArgs a) {
  constexpr int VEC = 16 / (int)sizeof(TI);
  constexpr int RX = MW / 2, RY = MH / 2;
  constexpr int HV = (RX + VEC - 1) / VEC;  // halo vectors on each side
  constexpr int NV = 64 + 2 * HV;           // vectors per staged row
  constexpr int LW = NV * VEC;
  constexpr int TW = 64 * VEC, TH = 4 * PPT;
  constexpr int LH = TH + 2 * RY;
  constexpr int NW = 1 + 2 * HV;            // window vectors per lane
  __shared__ __align__(16) TI tile[LH * LW];

  const Geom& g = a.g;
  const int tx0 = blockIdx.x * TW;
  const int ty0 = blockIdx.y * TH;
  const TI* in = static_cast<const TI*>(g.in) + (long long)blockIdx.z * g.in_plane;
  TO* out = static_cast<TO*>(g.out) + (long long)blockIdx.z * g.out_plane;

  stage_tile<TI, VEC, NV, LH>(tile, in, g, g.in_ox + tx0 - HV * VEC, g.in_oy + ty0 - RY,
                              a.in_vec_ok);
  __syncthreads();

  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  ACC acc[PPT][VEC];
#pragma unroll
  for (int o = 0; o < PPT; ++o)
#pragma unroll
    for (int v = 0; v < VEC; ++v) acc[o][v] = ACC(0);

  const TI* base = tile + (wave * PPT) * LW + lane * VEC;
#pragma unroll
  for (int ir = 0; ir < PPT + 2 * RY; ++ir) {
    ACC win[NW * VEC];
    load_window<TI, ACC, VEC, NW>(win, base + ir * LW);
#pragma unroll
    for (int o = 0; o < PPT; ++o) {
      const int dy = ir - o;
      if (dy < 0 || dy >= MH) continue;
#pragma unroll
      for (int dx = 0; dx < MW; ++dx) {
        ACC w;
        if constexpr (WDEV)
          w = (ACC)((const HH_CONST float*)a.wdev)[dy * MW + dx];
        else if constexpr (std::is_same<ACC, float>::value)
          w = a.w.f[dy * MW + dx];
        else
          w = (ACC)a.w.i[dy * MW + dx];
#pragma unroll
        for (int v = 0; v < VEC; ++v)
          acc[o][v] = madd<FMA>(w, win[HV * VEC + v + dx - RX], acc[o][v]);
      }
    }
  }

  const int ox = tx0 + lane * VEC;
#pragma unroll
  for (int o = 0; o < PPT; ++o)
    store_row<TO, ACC, VEC>(out, g, a.out_vec_ok, ox, ty0 + wave * PPT + o, acc[o]);
}

template <typename TI, typename TO, typename ACC, int MW, int MH, int PPT, bool FMA, bool WDEV>
int launch_sum(const LocalArgs& a, hipStream_t s) {
  constexpr int VEC = 16 / (int)sizeof(TI);
  dim3 grid(ceil_div(a.g.is_w, 64 * VEC), ceil_div(a.g.is_h, 4 * PPT), a.g.batch);
  hipLaunchKernelGGL((local_sum_kernel<TI, TO, ACC, MW, MH, PPT, FMA, WDEV>), grid, dim3(256), 0,
                     s, a);
  return hip_status(hipGetLastError(), "local_sum_kernel launch");
}

}  // namespace hh
