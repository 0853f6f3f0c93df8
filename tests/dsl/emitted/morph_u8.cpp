// Emitted by hipacc_amd.emit from morph_u8.cpp for the MI355X runtime (hipacc_amd/runtime/hipacc_hip.hpp).
// Integer local operators on one uchar image: Sobel-x with an int mask over
// Domain(mask) into short (MIRROR), a 3x3 erode (REPEAT) and a 5x5 box sum
// into uint (CONSTANT).  Writes the three outputs to argv[1], one after the
// other.
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

// Kernel class SobelX: local operator SUM over mask mask (hipaccLaunchLocal)

// Kernel class Erode: local operator MIN over domain dom (hipaccLaunchLocal)

// Kernel class BoxSum: local operator SUM over domain dom (hipaccLaunchLocal)

static uint32_t pixel_hash(int x, int y) {
    uint32_t h = (uint32_t)x * 2654435761u + (uint32_t)y * 40503u;
    h ^= h >> 16;
    h *= 0x7feb352du;
    return h ^ (h >> 15);
}

HIPACC_CODEGEN int main(int argc, const char **argv) {
    hipaccInitHIP(0);
    const int width = 777;
    const int height = 301;
    const int sobel[3][3] = { { -1, 0, 1 }, { -2, 0, 2 }, { -1, 0, 1 } };
    uchar *input = new uchar[width * height];
    for (int y = 0; y < height; ++y)
        for (int x = 0; x < width; ++x)
            input[y * width + x] = (uchar)(pixel_hash(x, y) >> 24);

    HipaccImageHip<uchar> in = hipaccCreateMemory<uchar>(input, width, height);
    HipaccImageHip<short> grad = hipaccCreateMemory<short>(nullptr, width, height);
    HipaccImageHip<uchar> eroded = hipaccCreateMemory<uchar>(nullptr, width, height);
    HipaccImageHip<uint> box = hipaccCreateMemory<uint>(nullptr, width, height);

    HipaccMask<int> smask(3, 3, &sobel[0][0]);
    auto sdom = smask;
    sdom.domain_from_coef();
    // BoundaryCondition bmirror: MIRROR of in
    HipaccAccessor<uchar> amirror = hipaccMakeAccessor(in);
    HipaccAccessor<short> igrad = hipaccMakeAccessor(grad);
    // SobelX sobel_x(igrad, amirror, sdom, smask): launched at sobel_x.execute()
    {
        HipaccLocalOptions sobel_x_opt;
        sobel_x_opt.reduce = HipaccReduce::SUM;
        sobel_x_opt.boundary = HipaccBoundary::MIRROR;
        sobel_x_opt.use_weights = true;
        sobel_x_opt.fp_contract = false;
        auto sobel_x_mask = hipaccMaskCast<int>(smask);
        sobel_x_mask.domain = sdom.domain;
        hipaccLaunchLocal(amirror, igrad, sobel_x_mask, sobel_x_opt);
    }

    HipaccMask<float> d3(3, 3);
    // BoundaryCondition brepeat: REPEAT of in
    HipaccAccessor<uchar> arepeat = hipaccMakeAccessor(in);
    HipaccAccessor<uchar> ieroded = hipaccMakeAccessor(eroded);
    // Erode erode(ieroded, arepeat, d3): launched at erode.execute()
    {
        HipaccLocalOptions erode_opt;
        erode_opt.reduce = HipaccReduce::MIN;
        erode_opt.boundary = HipaccBoundary::REPEAT;
        erode_opt.use_weights = false;
        erode_opt.fp_contract = false;
        auto erode_mask = hipaccMaskCast<int>(d3);
        hipaccLaunchLocal(arepeat, ieroded, erode_mask, erode_opt);
    }

    HipaccMask<float> d5(5, 5);
    // BoundaryCondition bconst: CONSTANT of in
    HipaccAccessor<uchar> aconst = hipaccMakeAccessor(in);
    HipaccAccessor<uint> ibox = hipaccMakeAccessor(box);
    // BoxSum box_sum(ibox, aconst, d5): launched at box_sum.execute()
    {
        HipaccLocalOptions box_sum_opt;
        box_sum_opt.reduce = HipaccReduce::SUM;
        box_sum_opt.boundary = HipaccBoundary::CONSTANT;
        box_sum_opt.const_val = 0;
        box_sum_opt.use_weights = false;
        box_sum_opt.fp_contract = false;
        auto box_sum_mask = hipaccMaskCast<int>(d5);
        hipaccLaunchLocal(aconst, ibox, box_sum_mask, box_sum_opt);
    }

    FILE *f = std::fopen(argv[1], "wb");
    std::fwrite(hipaccReadMemory<short>(grad), sizeof(short), (size_t)width * height, f);
    std::fwrite(hipaccReadMemory<uchar>(eroded), 1, (size_t)width * height, f);
    std::fwrite(hipaccReadMemory<uint>(box), sizeof(uint), (size_t)width * height, f);
    std::fclose(f);
    delete[] input;
    return 0;
}
