// Emitted by hipacc_amd.emit from reduce_f32.cpp for the MI355X runtime (hipacc_amd/runtime/hipacc_hip.hpp).
// Global reductions of a float image (Kernel::reduce and reduced_data(), in
// the form of the Reduction samples): minimum, maximum and sum.  Writes the
// three results to argv[1], then the sum and the sum of magnitudes in double
// and the pixel count.
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
using namespace hipacc::math;

// Kernel class MinReduction: point operator COPY (hipaccLaunchPoint); global MIN reduction (hipaccApplyReduction)

// Kernel class MaxReduction: point operator COPY (hipaccLaunchPoint); global MAX reduction (hipaccApplyReduction)

// Kernel class SumReduction: point operator COPY (hipaccLaunchPoint); global SUM reduction (hipaccApplyReduction)

static float pixel_value(int x, int y) {
    uint32_t h = (uint32_t)x * 2246822519u + (uint32_t)y * 3266489917u;
    h = (h ^ (h >> 15)) * 668265263u;
    return (float)((h ^ (h >> 13)) >> 8) * (1.0f / 16777216.0f) - 0.25f;
}

HIPACC_CODEGEN int main(int argc, const char **argv) {
    hipaccInitHIP(0);
    const int width = 1024;
    const int height = 333;
    float *input = new float[width * height];
    for (int y = 0; y < height; ++y)
        for (int x = 0; x < width; ++x)
            input[y * width + x] = pixel_value(x, y);

    HipaccImageHip<float> in = hipaccCreateMemory<float>(input, width, height);
    HipaccImageHip<float> out = hipaccCreateMemory<float>(nullptr, width, height);
    HipaccAccessor<float> acc = hipaccMakeAccessor(in);
    HipaccAccessor<float> iter = hipaccMakeAccessor(out);

    // MinReduction rmin(iter, acc): launched at rmin.execute()
    hipaccLaunchPoint(HIPACC_POINT_COPY, acc, iter);
    float vmin = hipaccApplyReduction<float>(iter, HipaccReduce::MIN);
    // MaxReduction rmax(iter, acc): launched at rmax.execute()
    hipaccLaunchPoint(HIPACC_POINT_COPY, acc, iter);
    float vmax = hipaccApplyReduction<float>(iter, HipaccReduce::MAX);
    // SumReduction rsum(iter, acc): launched at rsum.execute()
    hipaccLaunchPoint(HIPACC_POINT_COPY, acc, iter);
    float vsum = hipaccApplyReduction<float>(iter, HipaccReduce::SUM);

    // the sum in double and the sum of magnitudes, for the error bound of
    // the float sums
    double exact = 0.0, magnitude = 0.0;
    for (int i = 0; i < width * height; ++i) {
        exact += input[i];
        magnitude += input[i] < 0 ? -input[i] : input[i];
    }
    float result[3] = { vmin, vmax, vsum };
    double sums[3] = { exact, magnitude, (double)width * height };
    FILE *f = std::fopen(argv[1], "wb");
    std::fwrite(result, sizeof(float), 3, f);
    std::fwrite(sums, sizeof(double), 3, f);
    std::fclose(f);
    delete[] input;
    return 0;
}
