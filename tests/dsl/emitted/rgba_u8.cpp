// Emitted by hipacc_amd.emit from rgba_u8.cpp for the MI355X runtime (hipacc_amd/runtime/hipacc_hip.hpp).
// Kernels on 4-wide uchar4 pixels, in the form of the RGBA samples and of
// the motion-interpolation sample's vector binning: a float4 weighted sum
// through convert_float4 / convert_uchar4 (MIRROR), a componentwise MIN
// over a domain (CLAMP), an int4 contrast stretch with vector-scalar
// arithmetic and min / max, and per-tile float4 bins (sums of the channels
// and a count) combined by a reduce() of float4 values.  Every bin sum is an
// integer below 2^24, so it is exact in any order.  Writes the three images
// and the bins to argv[1].
#include "hipacc_hip_device.hpp"
// DSL names kept by the emitted host code
typedef unsigned char uchar;
typedef unsigned short ushort;
typedef unsigned int uint;
#ifndef HIPACC_CODEGEN
#define HIPACC_CODEGEN
#endif
namespace hipacc { namespace math { using std::min; using std::max; } }
namespace hipacc { using Reduce = hipacc_dev::Reduce; }

#include <cstdint>
#include <cstdio>

using namespace hipacc;
using namespace hipacc::math;

// Kernel class Smooth: device code, the kernel() body compiled as written (hipaccLaunchUser)
struct Smooth : hipacc_dev::Kernel<uchar4> {
    hipacc_dev::Accessor<uchar4> in;
    hipacc_dev::Mask<float> mask;
    __device__ void set_pos(int x_, int y_) { it_x = x_; it_y = y_; in.px = x_; in.py = y_; }
    __device__ __forceinline__ void kernel() {
        float4 s = convolve(mask, Reduce::SUM, [&] () -> float4 {
            return mask() * convert_float4(in(mask));
        });
        output() = convert_uchar4(s + 0.5f);
    }
};

// Kernel class Shrink: device code, the kernel() body compiled as written (hipaccLaunchUser)
struct Shrink : hipacc_dev::Kernel<uchar4> {
    hipacc_dev::Accessor<uchar4> in;
    hipacc_dev::Domain dom;
    __device__ void set_pos(int x_, int y_) { it_x = x_; it_y = y_; in.px = x_; in.py = y_; }
    __device__ __forceinline__ void kernel() {
        output() = reduce(dom, Reduce::MIN, [&] () -> uchar4 { return in(dom); });
    }
};

// Kernel class Stretch: device code, the kernel() body compiled as written (hipaccLaunchUser)
struct Stretch : hipacc_dev::Kernel<uchar4> {
    hipacc_dev::Accessor<uchar4> in;
    __device__ void set_pos(int x_, int y_) { it_x = x_; it_y = y_; in.px = x_; in.py = y_; }
    __device__ __forceinline__ void kernel() {
        int4 d = convert_int4(in()) - 128;
        output() = convert_uchar4(max(min(d * 3 + 128, 255), 0));
    }
};

// Kernel class TileSums: point operator COPY (hipaccLaunchPoint); binning() and reduce() as device code (hipaccBinUser)
struct TileSums_bins : hipacc_dev::Binning<float4> {
    int tiles_x;
    __device__ void binning(unsigned int x, unsigned int y, uchar4 pixel) {
        float4 v = convert_float4(pixel);
        v.w = 1.0f;
        bin((y / 16) * tiles_x + x / 16) = v;
    }
    __device__ float4 reduce(float4 left, float4 right) const { return left + right; }
};

static uint32_t mix(int x, int y, uint32_t salt) {
    uint32_t h = (uint32_t)x * 2654435761u ^ (uint32_t)y * 2246822519u ^ salt;
    h = (h ^ (h >> 16)) * 2146121005u;
    return h ^ (h >> 15);
}

HIPACC_CODEGEN int main(int argc, const char **argv) {
    hipaccInitHIP(0);
    const int width = 197;
    const int height = 131;
    const int tiles_x = (width + 15) / 16;
    const int tiles_y = (height + 15) / 16;
    const float coef[3][5] = {
        { 0.02f, 0.05f, 0.08f, 0.05f, 0.02f },
        { 0.05f, 0.10f, 0.16f, 0.10f, 0.05f },
        { 0.02f, 0.05f, 0.08f, 0.05f, 0.02f }
    };
    const uchar cross[3][3] = {
        { 0, 1, 0 },
        { 1, 1, 1 },
        { 0, 1, 0 }
    };
    uchar4 *pixels = new uchar4[width * height];
    for (int y = 0; y < height; ++y)
        for (int x = 0; x < width; ++x) {
            uint32_t h = mix(x, y, 3u);
            pixels[y * width + x] = make_uchar4((uchar)h, (uchar)(h >> 8), (uchar)(h >> 16),
                                                (uchar)(h >> 24));
        }

    HipaccImageHip<uchar4> in = hipaccCreateMemory<uchar4>(pixels, width, height);
    HipaccImageHip<uchar4> smooth = hipaccCreateMemory<uchar4>(nullptr, width, height);
    HipaccImageHip<uchar4> shrink = hipaccCreateMemory<uchar4>(nullptr, width, height);
    HipaccImageHip<uchar4> stretch = hipaccCreateMemory<uchar4>(nullptr, width, height);
    HipaccImageHip<uchar4> copy = hipaccCreateMemory<uchar4>(nullptr, width, height);
    HipaccMask<float> mask(5, 3, &coef[0][0]);
    HipaccMask<float> dom(3, 3);
    dom.domain_write(&cross[0][0]);

    // BoundaryCondition bc_mirror: MIRROR of in
    HipaccAccessor<uchar4> a_mirror = hipaccMakeAccessor(in);
    HipaccAccessor<uchar4> i_smooth = hipaccMakeAccessor(smooth);
    Smooth k_smooth;
    k_smooth.in = hipaccDevAccessor(a_mirror, HipaccBoundary::MIRROR, 0);
    k_smooth.mask = hipaccDevMask(mask);
    hipaccLaunchUser(k_smooth, i_smooth);

    // BoundaryCondition bc_clamp: CLAMP of in
    HipaccAccessor<uchar4> a_clamp = hipaccMakeAccessor(in);
    HipaccAccessor<uchar4> i_shrink = hipaccMakeAccessor(shrink);
    Shrink k_shrink;
    k_shrink.in = hipaccDevAccessor(a_clamp, HipaccBoundary::CLAMP, 0);
    k_shrink.dom = hipaccDevDomain(dom);
    hipaccLaunchUser(k_shrink, i_shrink);

    HipaccAccessor<uchar4> a_in = hipaccMakeAccessor(in);
    HipaccAccessor<uchar4> i_stretch = hipaccMakeAccessor(stretch);
    Stretch k_stretch;
    k_stretch.in = hipaccDevAccessor(a_in, HipaccBoundary::UNDEFINED, 0);
    hipaccLaunchUser(k_stretch, i_stretch);

    HipaccAccessor<uchar4> i_copy = hipaccMakeAccessor(copy);
    // TileSums k_tiles(i_copy, a_in, tiles_x): launched at k_tiles.execute()
    float4 *bins = [&] { hipaccLaunchPoint(HIPACC_POINT_COPY, a_in, i_copy); return hipaccBinUser<TileSums_bins, float4, HIPACC_REDUCE_SUM>([&] { TileSums_bins f; f.tiles_x = tiles_x; return f; }(), i_copy, tiles_x * tiles_y); }();

    FILE *f = std::fopen(argv[1], "wb");
    std::fwrite(hipaccReadMemory<uchar4>(smooth), sizeof(uchar4), width * height, f);
    std::fwrite(hipaccReadMemory<uchar4>(shrink), sizeof(uchar4), width * height, f);
    std::fwrite(hipaccReadMemory<uchar4>(stretch), sizeof(uchar4), width * height, f);
    std::fwrite(bins, sizeof(float4), tiles_x * tiles_y, f);
    std::fclose(f);
    delete[] bins;
    delete[] pixels;
    return 0;
}
