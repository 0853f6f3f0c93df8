// Emitted by hipacc_amd.emit from bilateral_u8.cpp for the MI355X runtime (hipacc_amd/runtime/hipacc_hip.hpp).
// The bilateral filter in the form of the Bilateral_Filter sample: a
// closeness mask computed on the host, a Domain over it, and iterate() with
// the range weight expf(-c_r * diff^2) per tap (CLAMP).  Writes the output
// image to argv[1].
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

#include <cmath>
#include <cstdint>
#include <cstdio>

using namespace hipacc;

// Kernel class BilateralFilter: device code, the kernel() body compiled as written (hipaccLaunchUser)
struct BilateralFilter : hipacc_dev::Kernel<uchar> {
    hipacc_dev::Accessor<uchar> in;
    hipacc_dev::Mask<float> mask;
    hipacc_dev::Domain dom;
    int sigma_r;
    __device__ void set_pos(int x_, int y_) { it_x = x_; it_y = y_; in.px = x_; in.py = y_; }
    __device__ __forceinline__ void kernel() {
        float c_r = 0.5f / (sigma_r * sigma_r);
        float d = 0.0f;
        float p = 0.0f;
        float center = in();

        iterate(dom, [&] () -> void {
            float diff = (float)(in(dom)) - center;
            float s = expf(-c_r * diff * diff) * mask(dom);
            d += s;
            p += s * in(dom);
        });

        output() = (uchar)(p / d + 0.5f);
    }
};

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
    BilateralFilter filter;
    filter.in = hipaccDevAccessor(acc, HipaccBoundary::CLAMP, 0);
    filter.mask = hipaccDevMask(mask);
    filter.dom = hipaccDevDomain(dom);
    filter.sigma_r = sigma_r;
    hipaccLaunchUser(filter, iter);

    FILE *f = std::fopen(argv[1], "wb");
    std::fwrite(hipaccReadMemory<uchar>(out), 1, (size_t)width * height, f);
    std::fclose(f);
    delete[] input;
    return 0;
}
