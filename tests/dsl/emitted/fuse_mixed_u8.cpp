// Emitted by hipacc_amd.emit from fuse_mixed_u8.cpp for the MI355X runtime (hipacc_amd/runtime/hipacc_hip.hpp).
// The kernel chain of the Mixed fusion sample on uchar images: point, point,
// 5x5 local, point, point, 5x5 local, point, point, each kernel over the
// previous one's image (point: v = in(); v += 3; local: the rounded 5x5
// Gaussian, CLAMP).  The intermediate images are named only by the chain,
// so the emitter launches the chain as one fused kernel.  Writes the last
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

#include <cstdint>
#include <cstdio>

using namespace hipacc;

// Kernel class PointOperator: point operator AFFINE (hipaccLaunchPoint)

// Kernel class LocalOperator: local operator SUM over mask mask, rounded (hipaccLaunchLocal)

static uint32_t pixel_hash(int x, int y) {
    uint32_t h = (uint32_t)x * 2654435761u ^ (uint32_t)y * 97u;
    h ^= h >> 15;
    h *= 0x2c1b3c6du;
    return h ^ (h >> 12);
}

HIPACC_CODEGEN int main(int argc, const char **argv) {
    hipaccInitHIP(0);
    const int width = 523;
    const int height = 317;
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
            input[y * width + x] = (uchar)(pixel_hash(x, y) >> 24);

    HipaccImageHip<uchar> in = hipaccCreateMemory<uchar>(input, width, height);
    HipaccImageHip<uchar> out = hipaccCreateMemory<uchar>(nullptr, width, height);
    HipaccMask<float> mask(5, 5, &coef[0][0]);

    HipaccImageHip<uchar> buf0 = hipaccCreateMemory<uchar>(nullptr, width, height);
    HipaccAccessor<uchar> acc0 = hipaccMakeAccessor(in);
    HipaccAccessor<uchar> iter0 = hipaccMakeAccessor(buf0);
    // PointOperator pointOp0(iter0, acc0): launched at pointOp0.execute()

    HipaccImageHip<uchar> buf1 = hipaccCreateMemory<uchar>(nullptr, width, height);
    HipaccAccessor<uchar> acc1 = hipaccMakeAccessor(buf0);
    HipaccAccessor<uchar> iter1 = hipaccMakeAccessor(buf1);
    // PointOperator pointOp1(iter1, acc1): launched at pointOp1.execute()

    HipaccImageHip<uchar> buf2 = hipaccCreateMemory<uchar>(nullptr, width, height);
    // BoundaryCondition bound2: CLAMP of buf1
    HipaccAccessor<uchar> acc2 = hipaccMakeAccessor(buf1);
    HipaccAccessor<uchar> iter2 = hipaccMakeAccessor(buf2);
    // LocalOperator localOp2(iter2, acc2, mask): launched at localOp2.execute()

    HipaccImageHip<uchar> buf3 = hipaccCreateMemory<uchar>(nullptr, width, height);
    HipaccAccessor<uchar> acc3 = hipaccMakeAccessor(buf2);
    HipaccAccessor<uchar> iter3 = hipaccMakeAccessor(buf3);
    // PointOperator pointOp3(iter3, acc3): launched at pointOp3.execute()

    HipaccImageHip<uchar> buf4 = hipaccCreateMemory<uchar>(nullptr, width, height);
    HipaccAccessor<uchar> acc4 = hipaccMakeAccessor(buf3);
    HipaccAccessor<uchar> iter4 = hipaccMakeAccessor(buf4);
    // PointOperator pointOp4(iter4, acc4): launched at pointOp4.execute()

    HipaccImageHip<uchar> buf5 = hipaccCreateMemory<uchar>(nullptr, width, height);
    // BoundaryCondition bound5: CLAMP of buf4
    HipaccAccessor<uchar> acc5 = hipaccMakeAccessor(buf4);
    HipaccAccessor<uchar> iter5 = hipaccMakeAccessor(buf5);
    // LocalOperator localOp5(iter5, acc5, mask): launched at localOp5.execute()

    HipaccImageHip<uchar> buf6 = hipaccCreateMemory<uchar>(nullptr, width, height);
    HipaccAccessor<uchar> acc6 = hipaccMakeAccessor(buf5);
    HipaccAccessor<uchar> iter6 = hipaccMakeAccessor(buf6);
    // PointOperator pointOp6(iter6, acc6): launched at pointOp6.execute()

    HipaccAccessor<uchar> acc7 = hipaccMakeAccessor(buf6);
    HipaccAccessor<uchar> iter7 = hipaccMakeAccessor(out);
    // PointOperator pointOp7(iter7, acc7): launched at pointOp7.execute()

    // pointOp0.execute(): fused into the launch of pointOp7
    // pointOp1.execute(): fused into the launch of pointOp7
    // localOp2.execute(): fused into the launch of pointOp7
    // pointOp3.execute(): fused into the launch of pointOp7
    // pointOp4.execute(): fused into the launch of pointOp7
    // localOp5.execute(): fused into the launch of pointOp7
    // pointOp6.execute(): fused into the launch of pointOp7
    {
        // fused: pointOp0 -> pointOp1 -> localOp2 -> pointOp3 -> pointOp4 -> localOp5 -> pointOp6 -> pointOp7
        std::vector<HipaccFusedStage> pointOp7_chain;
        pointOp7_chain.push_back(hipaccFusedPoint<uchar>(HIPACC_POINT_AFFINE, 1.0f, 3));
        pointOp7_chain.push_back(hipaccFusedPoint<uchar>(HIPACC_POINT_AFFINE, 1.0f, 3));
        HipaccLocalOptions localOp2_opt;
        localOp2_opt.reduce = HipaccReduce::SUM;
        localOp2_opt.boundary = HipaccBoundary::CLAMP;
        localOp2_opt.use_weights = true;
        localOp2_opt.round = true;
        localOp2_opt.fp_contract = false;
        auto localOp2_mask = hipaccMaskCast<float>(mask);
        pointOp7_chain.push_back(hipaccFusedLocal<uchar>(localOp2_mask, localOp2_opt));
        pointOp7_chain.push_back(hipaccFusedPoint<uchar>(HIPACC_POINT_AFFINE, 1.0f, 3));
        pointOp7_chain.push_back(hipaccFusedPoint<uchar>(HIPACC_POINT_AFFINE, 1.0f, 3));
        HipaccLocalOptions localOp5_opt;
        localOp5_opt.reduce = HipaccReduce::SUM;
        localOp5_opt.boundary = HipaccBoundary::CLAMP;
        localOp5_opt.use_weights = true;
        localOp5_opt.round = true;
        localOp5_opt.fp_contract = false;
        auto localOp5_mask = hipaccMaskCast<float>(mask);
        pointOp7_chain.push_back(hipaccFusedLocal<uchar>(localOp5_mask, localOp5_opt));
        pointOp7_chain.push_back(hipaccFusedPoint<uchar>(HIPACC_POINT_AFFINE, 1.0f, 3));
        pointOp7_chain.push_back(hipaccFusedPoint<uchar>(HIPACC_POINT_AFFINE, 1.0f, 3));
        hipaccLaunchFused(acc0, iter7, pointOp7_chain);
    }

    FILE *f = std::fopen(argv[1], "wb");
    std::fwrite(hipaccReadMemory<uchar>(out), 1, (size_t)width * height, f);
    std::fclose(f);
    delete[] input;
    return 0;
}
