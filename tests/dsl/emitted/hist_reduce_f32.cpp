// Emitted by hipacc_amd.emit from hist_reduce_f32.cpp for the MI355X runtime (hipacc_amd/runtime/hipacc_hip.hpp).
// Global operators the library has no fixed operator for, in the form of the
// Histogram sample (Kernel<float, uint> with binning() and reduce()) and a
// reduction whose reduce() keeps the value of largest magnitude.  Writes the
// 64 bins and the reduction's value to argv[1].
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
using namespace hipacc::math;

// Kernel class Histogram: point operator COPY (hipaccLaunchPoint); binning() and reduce() as device code (hipaccBinUser)
struct Histogram_bins : hipacc_dev::Binning<uint> {
    __device__ void binning(unsigned int x, unsigned int y, float pixel) {
        bin(pixel / 255.0f * num_bins()) = 1;
    }
    __device__ uint reduce(uint left, uint right) const { return left + right; }
};

// Kernel class MaxMagnitude: point operator AFFINE (hipaccLaunchPoint); reduce() as device code (hipaccReduceUser)
struct MaxMagnitude_reduce {
    __device__ float reduce(float left, float right) const {
        return fabsf(left) > fabsf(right) ? left : right;
    }
};

static float pixel_value(int x, int y) {
    uint32_t h = (uint32_t)x * 3266489917u + (uint32_t)y * 668265263u;
    h = (h ^ (h >> 15)) * 2246822519u;
    return (float)((h ^ (h >> 13)) >> 8) * (255.0f / 16777216.0f);
}

HIPACC_CODEGEN int main(int argc, const char **argv) {
    hipaccInitHIP(0);
    const int width = 517;
    const int height = 263;
    const int num_bins = 64;
    float *input = new float[width * height];
    for (int y = 0; y < height; ++y)
        for (int x = 0; x < width; ++x)
            input[y * width + x] = pixel_value(x, y);

    HipaccImageHip<float> in = hipaccCreateMemory<float>(input, width, height);
    HipaccImageHip<float> copy = hipaccCreateMemory<float>(nullptr, width, height);
    HipaccImageHip<float> shifted = hipaccCreateMemory<float>(nullptr, width, height);

    HipaccAccessor<float> acc = hipaccMakeAccessor(in);
    HipaccAccessor<float> icopy = hipaccMakeAccessor(copy);
    // Histogram hist(icopy, acc): launched at hist.execute()
    hipaccLaunchPoint(HIPACC_POINT_COPY, acc, icopy);
    uint *bins = hipaccBinUser<Histogram_bins, uint, HIPACC_REDUCE_SUM>(Histogram_bins{}, icopy, num_bins);

    HipaccAccessor<float> ishifted = hipaccMakeAccessor(shifted);
    // MaxMagnitude maxmag(ishifted, acc): launched at maxmag.execute()
    hipaccLaunchPoint(HIPACC_POINT_AFFINE, acc, ishifted, static_cast<const HipaccAccessor<float>*>(nullptr), 1.0f, -100.0f);
    float extreme = hipaccReduceUser(MaxMagnitude_reduce{}, ishifted);

    FILE *f = std::fopen(argv[1], "wb");
    std::fwrite(bins, sizeof(uint), num_bins, f);
    std::fwrite(&extreme, sizeof(float), 1, f);
    std::fclose(f);
    delete[] bins;
    delete[] input;
    return 0;
}
