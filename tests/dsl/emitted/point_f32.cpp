// Emitted by hipacc_amd.emit from point_f32.cpp for the MI355X runtime (hipacc_amd/runtime/hipacc_hip.hpp).
// Point operators and a float local operator on float images: a scale and
// offset, the fusion samples' `v += 3` point kernel, a difference of two
// images and a 3x3 float convolution (MIRROR) of the difference.  Writes the
// outputs to argv[1], one after the other.
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

// Kernel class Scale: point operator AFFINE (hipaccLaunchPoint)

// Kernel class AddThree: point operator AFFINE (hipaccLaunchPoint)

// Kernel class Difference: point operator SUB (hipaccLaunchPoint)

// Kernel class Blur3: local operator SUM over mask mask (hipaccLaunchLocal)

static float pixel_value(int x, int y) {
    uint32_t h = (uint32_t)x * 374761393u + (uint32_t)y * 668265263u;
    h = (h ^ (h >> 13)) * 1274126177u;
    return (float)((h ^ (h >> 16)) >> 8) * (1.0f / 16777216.0f);
}

HIPACC_CODEGEN int main(int argc, const char **argv) {
    hipaccInitHIP(0);
    const int width = 640;
    const int height = 257;
    const float coef[3][3] = { { 0.0625f, 0.125f, 0.0625f },
                               { 0.125f, 0.25f, 0.125f },
                               { 0.0625f, 0.125f, 0.0625f } };
    float *input = new float[width * height];
    for (int y = 0; y < height; ++y)
        for (int x = 0; x < width; ++x)
            input[y * width + x] = pixel_value(x, y);

    HipaccImageHip<float> in = hipaccCreateMemory<float>(input, width, height);
    HipaccImageHip<float> scaled = hipaccCreateMemory<float>(nullptr, width, height);
    HipaccImageHip<float> shifted = hipaccCreateMemory<float>(nullptr, width, height);
    HipaccImageHip<float> diff = hipaccCreateMemory<float>(nullptr, width, height);
    HipaccImageHip<float> blurred = hipaccCreateMemory<float>(nullptr, width, height);

    HipaccAccessor<float> ain = hipaccMakeAccessor(in);
    HipaccAccessor<float> iscaled = hipaccMakeAccessor(scaled);
    // Scale scale(iscaled, ain): launched at scale.execute()
    hipaccLaunchPoint(HIPACC_POINT_AFFINE, ain, iscaled, static_cast<const HipaccAccessor<float>*>(nullptr), 0.5f, 0.25f);

    HipaccAccessor<float> ascaled = hipaccMakeAccessor(scaled);
    HipaccAccessor<float> ishifted = hipaccMakeAccessor(shifted);
    // AddThree add_three(ishifted, ascaled): launched at add_three.execute()
    hipaccLaunchPoint(HIPACC_POINT_AFFINE, ascaled, ishifted, static_cast<const HipaccAccessor<float>*>(nullptr), 1.0f, 3);

    HipaccAccessor<float> ashifted = hipaccMakeAccessor(shifted);
    HipaccAccessor<float> idiff = hipaccMakeAccessor(diff);
    // Difference difference(idiff, ashifted, ain): launched at difference.execute()
    hipaccLaunchPoint(HIPACC_POINT_SUB, ashifted, idiff, &ain);

    HipaccMask<float> mask(3, 3, &coef[0][0]);
    // BoundaryCondition bdiff: MIRROR of diff
    HipaccAccessor<float> adiff = hipaccMakeAccessor(diff);
    HipaccAccessor<float> iblurred = hipaccMakeAccessor(blurred);
    // Blur3 blur(iblurred, adiff, mask): launched at blur.execute()
    {
        HipaccLocalOptions blur_opt;
        blur_opt.reduce = HipaccReduce::SUM;
        blur_opt.boundary = HipaccBoundary::MIRROR;
        blur_opt.use_weights = true;
        blur_opt.fp_contract = false;
        auto blur_mask = hipaccMaskCast<float>(mask);
        hipaccLaunchLocal(adiff, iblurred, blur_mask, blur_opt);
    }

    FILE *f = std::fopen(argv[1], "wb");
    std::fwrite(hipaccReadMemory<float>(scaled), sizeof(float), (size_t)width * height, f);
    std::fwrite(hipaccReadMemory<float>(shifted), sizeof(float), (size_t)width * height, f);
    std::fwrite(hipaccReadMemory<float>(diff), sizeof(float), (size_t)width * height, f);
    std::fwrite(hipaccReadMemory<float>(blurred), sizeof(float), (size_t)width * height, f);
    std::fclose(f);
    delete[] input;
    return 0;
}
