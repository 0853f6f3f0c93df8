// Emitted by hipacc_amd.emit from bilateral_u8.cpp for the MI355X runtime (hipacc_amd/runtime/hipacc_hip.hpp).
// The bilateral filter in the form of the Bilateral_Filter sample: a
// closeness mask computed on the host, a Domain over it, and iterate() with
// the range weight expf(-c_r * diff^2) per tap (CLAMP).  Writes the output
// image to argv[1].
#include "hipacc_hip.hpp"
// DSL names kept by the emitted host code
typedef unsigned char uchar;
typedef unsigned short ushort;
typedef unsigned int uint;
#ifndef HIPACC_CODEGEN
#define HIPACC_CODEGEN
#endif
namespace hipacc { namespace math { using std::min; using std::max; } }

#include <cmath>
#include <cstdint>
#include <cstdio>

using namespace hipacc;

// Kernel class BilateralFilter: bilateral filter over domain dom, spatial weights mask, rounded (hipaccLaunchBilateral)

static uint32_t pixel_hash(int x, int y) {
    uint32_t h = (uint32_t)(x / 3) * 73856093u ^ (uint32_t)(y / 3) * 19349663u;
    h ^= h >> 13;
    h *= 0x5bd1e995u;
    return (h ^ (h >> 15)) + (uint32_t)(x + y);
}

HIPACC_CODEGEN int main(int argc, const char **argv) {
    hipaccInitHIP(0);
    const int width = 301;
    const int height = 199;
    const int sigma_d = 2;
    const int sigma_r = 16;
    float coef[7][7];
    for (int yf = -3; yf <= 3; ++yf)
        for (int xf = -3; xf <= 3; ++xf)
            coef[yf + 3][xf + 3] = expf(-0.5f / (sigma_d * sigma_d) * (xf * xf + yf * yf));
    uchar *input = new uchar[width * height];
    for (int y = 0; y < height; ++y)
        for (int x = 0; x < width; ++x)
            input[y * width + x] = (uchar)(pixel_hash(x, y) & 255);

    HipaccImageHip<uchar> in = hipaccCreateMemory<uchar>(input, width, height);
    HipaccImageHip<uchar> out = hipaccCreateMemory<uchar>(nullptr, width, height);
    HipaccMask<float> mask(7, 7, &coef[0][0]);
    auto dom = mask;
    dom.domain_from_coef();
    // BoundaryCondition bound: CLAMP of in
    HipaccAccessor<uchar> acc = hipaccMakeAccessor(in);
    HipaccAccessor<uchar> iter = hipaccMakeAccessor(out);
    // BilateralFilter filter(iter, acc, mask, dom, sigma_r): launched at filter.execute()
    {
        auto filter_spatial = hipaccBilateralSpatial(mask, dom);
        hipaccLaunchBilateral(acc, iter, mask.size_x, 0.0f, (float)(sigma_r), filter_spatial.data(), HipaccBoundary::CLAMP, 0, true);
    }

    FILE *f = std::fopen(argv[1], "wb");
    std::fwrite(hipaccReadMemory<uchar>(out), 1, (size_t)width * height, f);
    std::fclose(f);
    delete[] input;
    return 0;
}
