// Emitted by hipacc_amd.emit from unsharp_f32.cpp for the MI355X runtime (hipacc_amd/runtime/hipacc_hip.hpp).
// Kernels the library has no operator for, in the form of the Unsharp and
// Laplace samples: a per-pixel polynomial, a weighted difference of two
// images, a ratio, and a 5-point Laplacian written with explicit offsets
// (MIRROR); the 5x5 blur between them is a library convolution.  Writes the
// outputs to argv[1], one after the other.
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

// Kernel class Tone: device code, the kernel() body compiled as written (hipaccLaunchUser)
struct Tone : hipacc_dev::Kernel<float> {
    hipacc_dev::Accessor<float> input;
    __device__ void set_pos(int x_, int y_) { it_x = x_; it_y = y_; input.px = x_; input.py = y_; }
    __device__ __forceinline__ void kernel() {
        float v = input();
        output() = v * v * 0.5f + v * 0.5f;
    }
};

// Kernel class Blur: local operator SUM over mask mask (hipaccLaunchLocal)

// Kernel class Sharpen: device code, the kernel() body compiled as written (hipaccLaunchUser)
struct Sharpen : hipacc_dev::Kernel<float> {
    hipacc_dev::Accessor<float> Gray;
    hipacc_dev::Accessor<float> InputBlur;
    __device__ void set_pos(int x_, int y_) { it_x = x_; it_y = y_; Gray.px = x_; Gray.py = y_; InputBlur.px = x_; InputBlur.py = y_; }
    __device__ __forceinline__ void kernel() { output() = 2 * Gray() - InputBlur(); }
};

// Kernel class Ratio: device code, the kernel() body compiled as written (hipaccLaunchUser)
struct Ratio : hipacc_dev::Kernel<float> {
    hipacc_dev::Accessor<float> Sharp;
    hipacc_dev::Accessor<float> Gray;
    float eps;
    __device__ void set_pos(int x_, int y_) { it_x = x_; it_y = y_; Sharp.px = x_; Sharp.py = y_; Gray.px = x_; Gray.py = y_; }
    __device__ __forceinline__ void kernel() { output() = Sharp() / (Gray() + eps); }
};

// Kernel class Laplace: device code, the kernel() body compiled as written (hipaccLaunchUser)
struct Laplace : hipacc_dev::Kernel<float> {
    hipacc_dev::Accessor<float> input;
    __device__ void set_pos(int x_, int y_) { it_x = x_; it_y = y_; input.px = x_; input.py = y_; }
    __device__ __forceinline__ void kernel() {
        output() = 4.0f * input() - input(-1, 0) - input(1, 0) - input(0, -1) - input(0, 1)
                   + 0.25f * (float)(x() % 7) - 0.125f * (float)(y() % 5);
    }
};

static float pixel_value(int x, int y) {
    uint32_t h = (uint32_t)x * 1597334677u + (uint32_t)y * 3812015801u;
    h = (h ^ (h >> 16)) * 2246822519u;
    return (float)((h ^ (h >> 13)) >> 8) * (1.0f / 16777216.0f);
}

HIPACC_CODEGEN int main(int argc, const char **argv) {
    hipaccInitHIP(0);
    const int width = 403;
    const int height = 211;
    const float coef[5][5] = {
        { 0.005008f, 0.017300f, 0.026151f, 0.017300f, 0.005008f },
        { 0.017300f, 0.059761f, 0.090339f, 0.059761f, 0.017300f },
        { 0.026151f, 0.090339f, 0.136565f, 0.090339f, 0.026151f },
        { 0.017300f, 0.059761f, 0.090339f, 0.059761f, 0.017300f },
        { 0.005008f, 0.017300f, 0.026151f, 0.017300f, 0.005008f }
    };
    float *input = new float[width * height];
    for (int y = 0; y < height; ++y)
        for (int x = 0; x < width; ++x)
            input[y * width + x] = pixel_value(x, y);

    HipaccImageHip<float> in = hipaccCreateMemory<float>(input, width, height);
    HipaccImageHip<float> gray = hipaccCreateMemory<float>(nullptr, width, height);
    HipaccImageHip<float> blurred = hipaccCreateMemory<float>(nullptr, width, height);
    HipaccImageHip<float> sharp = hipaccCreateMemory<float>(nullptr, width, height);
    HipaccImageHip<float> ratio = hipaccCreateMemory<float>(nullptr, width, height);
    HipaccImageHip<float> lap = hipaccCreateMemory<float>(nullptr, width, height);

    HipaccAccessor<float> ain = hipaccMakeAccessor(in);
    HipaccAccessor<float> igray = hipaccMakeAccessor(gray);
    Tone tone;
    tone.input = hipaccDevAccessor(ain, HipaccBoundary::UNDEFINED, 0);
    hipaccLaunchUser(tone, igray);

    HipaccMask<float> mask(5, 5, &coef[0][0]);
    // BoundaryCondition bgray: CLAMP of gray
    HipaccAccessor<float> agray_bc = hipaccMakeAccessor(gray);
    HipaccAccessor<float> iblurred = hipaccMakeAccessor(blurred);
    // Blur blur(iblurred, agray_bc, mask): launched at blur.execute()
    {
        HipaccLocalOptions blur_opt;
        blur_opt.reduce = HipaccReduce::SUM;
        blur_opt.boundary = HipaccBoundary::CLAMP;
        blur_opt.use_weights = true;
        blur_opt.fp_contract = false;
        auto blur_mask = hipaccMaskCast<float>(mask);
        hipaccLaunchLocal(agray_bc, iblurred, blur_mask, blur_opt);
    }

    HipaccAccessor<float> agray = hipaccMakeAccessor(gray);
    HipaccAccessor<float> ablurred = hipaccMakeAccessor(blurred);
    HipaccAccessor<float> isharp = hipaccMakeAccessor(sharp);
    Sharpen sharpen;
    sharpen.Gray = hipaccDevAccessor(agray, HipaccBoundary::UNDEFINED, 0);
    sharpen.InputBlur = hipaccDevAccessor(ablurred, HipaccBoundary::UNDEFINED, 0);
    hipaccLaunchUser(sharpen, isharp);

    HipaccAccessor<float> asharp = hipaccMakeAccessor(sharp);
    HipaccAccessor<float> iratio = hipaccMakeAccessor(ratio);
    Ratio rat;
    rat.Sharp = hipaccDevAccessor(asharp, HipaccBoundary::UNDEFINED, 0);
    rat.Gray = hipaccDevAccessor(agray, HipaccBoundary::UNDEFINED, 0);
    rat.eps = 0.01f;
    hipaccLaunchUser(rat, iratio);

    // BoundaryCondition bmirror: MIRROR of in
    HipaccAccessor<float> amirror = hipaccMakeAccessor(in);
    HipaccAccessor<float> ilap = hipaccMakeAccessor(lap);
    Laplace laplace;
    laplace.input = hipaccDevAccessor(amirror, HipaccBoundary::MIRROR, 0);
    hipaccLaunchUser(laplace, ilap);

    FILE *f = std::fopen(argv[1], "wb");
    std::fwrite(hipaccReadMemory<float>(gray), sizeof(float), (size_t)width * height, f);
    std::fwrite(hipaccReadMemory<float>(blurred), sizeof(float), (size_t)width * height, f);
    std::fwrite(hipaccReadMemory<float>(sharp), sizeof(float), (size_t)width * height, f);
    std::fwrite(hipaccReadMemory<float>(ratio), sizeof(float), (size_t)width * height, f);
    std::fwrite(hipaccReadMemory<float>(lap), sizeof(float), (size_t)width * height, f);
    std::fclose(f);
    delete[] input;
    return 0;
}
