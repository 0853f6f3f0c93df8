// Emitted by hipacc_amd.emit from gauss_u8.cpp for the MI355X runtime (hipacc_amd/runtime/hipacc_hip.hpp).
// Gaussian blur of a uchar image: 5x5 mask, CLAMP, (uchar)(sum + 0.5f), in
// the form of the Gaussian_Blur sample.  Writes the output image to argv[1].
#include "hipacc_hip.hpp"
// DSL names kept by the emitted host code
typedef unsigned char uchar;
typedef unsigned short ushort;
typedef unsigned int uint;
#ifndef HIPACC_CODEGEN
#define HIPACC_CODEGEN
#endif
namespace hipacc { namespace math { using std::min; using std::max; } }

#include <cstdint>
#include <cstdio>

using namespace hipacc;

// Kernel class GaussianBlur: local operator SUM over mask mask, rounded (hipaccLaunchLocal)

static uint32_t pixel_hash(int x, int y) {
    uint32_t h = (uint32_t)x * 73856093u ^ (uint32_t)y * 19349663u;
    h ^= h >> 13;
    h *= 0x5bd1e995u;
    return h ^ (h >> 15);
}

HIPACC_CODEGEN int main(int argc, const char **argv) {
    hipaccInitHIP(0);
    const int width = 1000;
    const int height = 600;
    const float coef[5][5] = {
        { 0.005008f, 0.017300f, 0.026151f, 0.017300f, 0.005008f },
        { 0.017300f, 0.059761f, 0.090339f, 0.059761f, 0.017300f },
        { 0.026151f, 0.090339f, 0.136565f, 0.090339f, 0.026151f },
        { 0.017300f, 0.059761f, 0.090339f, 0.059761f, 0.017300f },
        { 0.005008f, 0.017300f, 0.026151f, 0.017300f, 0.005008f }
    };
    uchar *input = new uchar[width * height];
    for (int y = 0; y < height; ++y)
        for (int x = 0; x < width; ++x)
            input[y * width + x] = (uchar)(pixel_hash(x, y) & 255);

    HipaccImageHip<uchar> in = hipaccCreateMemory<uchar>(input, width, height);
    HipaccImageHip<uchar> out = hipaccCreateMemory<uchar>(nullptr, width, height);
    HipaccMask<float> mask(5, 5, &coef[0][0]);
    // BoundaryCondition bound: CLAMP of in
    HipaccAccessor<uchar> acc = hipaccMakeAccessor(in);
    HipaccAccessor<uchar> iter = hipaccMakeAccessor(out);
    // GaussianBlur filter(iter, acc, mask): launched at filter.execute()
    {
        HipaccLocalOptions filter_opt;
        filter_opt.reduce = HipaccReduce::SUM;
        filter_opt.boundary = HipaccBoundary::CLAMP;
        filter_opt.use_weights = true;
        filter_opt.round = true;
        filter_opt.fp_contract = false;
        auto filter_mask = hipaccMaskCast<float>(mask);
        hipaccLaunchLocal(acc, iter, filter_mask, filter_opt);
    }
    uchar *output = hipaccReadMemory<uchar>(out);

    FILE *f = std::fopen(argv[1], "wb");
    std::fwrite(output, 1, (size_t)width * height, f);
    std::fclose(f);
    delete[] input;
    return 0;
}
