// Emitted by hipacc_amd.emit from interp_f32.cpp for the MI355X runtime (hipacc_amd/runtime/hipacc_hip.hpp).
// Interpolating accessors read by kernels that are not library operators:
// a float image scaled up (x 1.7, x 1.3) and down (x 0.6) through accessors
// with Interpolate::NN, LF, B5, CF and L3, each over a boundary condition
// (CLAMP, MIRROR, REPEAT, CONSTANT), read at the pixel and at an offset,
// with the result scaled and offset in the kernel.  Writes the up- and
// down-scaled images of every interpolation mode to argv[1].
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

// Kernel class Resize: device code, the kernel() body compiled as written (hipaccLaunchUser)
struct Resize : hipacc_dev::Kernel<float> {
    hipacc_dev::Accessor<float, true> in;
    float gain;
    __device__ void set_pos(int x_, int y_) { it_x = x_; it_y = y_; in.px = x_; in.py = y_; }
    __device__ __forceinline__ void kernel() { output() = gain * in() + 0.25f * in(1, -1); }
};

static float value(int x, int y) {
    uint32_t h = (uint32_t)x * 2654435761u ^ (uint32_t)y * 2246822519u;
    h = (h ^ (h >> 16)) * 2146121005u;
    return (float)((h ^ (h >> 15)) >> 8) * (1.0f / 16777216.0f) - 0.5f;
}

HIPACC_CODEGEN int main(int argc, const char **argv) {
    hipaccInitHIP(0);
    const int width = 123;
    const int height = 77;
    const int up_w = 209, up_h = 100;     // x 1.7, x 1.3
    const int down_w = 74, down_h = 46;   // x 0.6
    const float dummy[3][3] = { { 0.f, 0.f, 0.f }, { 0.f, 1.f, 0.f }, { 0.f, 0.f, 0.f } };
    float *pixels = new float[width * height];
    for (int y = 0; y < height; ++y)
        for (int x = 0; x < width; ++x)
            pixels[y * width + x] = value(x, y);

    HipaccImageHip<float> in = hipaccCreateMemory<float>(pixels, width, height);
    HipaccMask<float> mask(3, 3, &dummy[0][0]);
    HipaccImageHip<float> up_nn = hipaccCreateMemory<float>(nullptr, up_w, up_h);
    HipaccImageHip<float> up_lf = hipaccCreateMemory<float>(nullptr, up_w, up_h);
    HipaccImageHip<float> up_b5 = hipaccCreateMemory<float>(nullptr, up_w, up_h);
    HipaccImageHip<float> up_cf = hipaccCreateMemory<float>(nullptr, up_w, up_h);
    HipaccImageHip<float> up_l3 = hipaccCreateMemory<float>(nullptr, up_w, up_h);
    HipaccImageHip<float> dn_nn = hipaccCreateMemory<float>(nullptr, down_w, down_h);
    HipaccImageHip<float> dn_lf = hipaccCreateMemory<float>(nullptr, down_w, down_h);
    HipaccImageHip<float> dn_b5 = hipaccCreateMemory<float>(nullptr, down_w, down_h);
    HipaccImageHip<float> dn_cf = hipaccCreateMemory<float>(nullptr, down_w, down_h);
    HipaccImageHip<float> dn_l3 = hipaccCreateMemory<float>(nullptr, down_w, down_h);

    // BoundaryCondition b_clamp: CLAMP of in
    // BoundaryCondition b_mirror: MIRROR of in
    // BoundaryCondition b_repeat: REPEAT of in
    // the DSL host emulation reads 0 whatever constant is given (its
    // BoundaryCondition never stores the value, dsl/image.hpp:259-290), so
    // the constant here is 0
    // BoundaryCondition b_const: CONSTANT of in

    HipaccAccessor<float> a_nn = hipaccMakeAccessor(in);
    HipaccAccessor<float> a_lf = hipaccMakeAccessor(in);
    HipaccAccessor<float> a_b5 = hipaccMakeAccessor(in);
    HipaccAccessor<float> a_cf = hipaccMakeAccessor(in);
    HipaccAccessor<float> a_l3 = hipaccMakeAccessor(in);

    HipaccAccessor<float> i_up_nn = hipaccMakeAccessor(up_nn);
    HipaccAccessor<float> i_up_lf = hipaccMakeAccessor(up_lf);
    HipaccAccessor<float> i_up_b5 = hipaccMakeAccessor(up_b5);
    HipaccAccessor<float> i_up_cf = hipaccMakeAccessor(up_cf);
    HipaccAccessor<float> i_up_l3 = hipaccMakeAccessor(up_l3);
    Resize r_up_nn;
    r_up_nn.in = hipaccDevAccessor<true>(a_nn, HipaccBoundary::CLAMP, 0, HIPACC_INTERP_NN, &i_up_nn);
    r_up_nn.gain = 1.0f;
    hipaccLaunchUser(r_up_nn, i_up_nn);
    Resize r_up_lf;
    r_up_lf.in = hipaccDevAccessor<true>(a_lf, HipaccBoundary::MIRROR, 0, HIPACC_INTERP_LF, &i_up_lf);
    r_up_lf.gain = 0.5f;
    hipaccLaunchUser(r_up_lf, i_up_lf);
    Resize r_up_b5;
    r_up_b5.in = hipaccDevAccessor<true>(a_b5, HipaccBoundary::REPEAT, 0, HIPACC_INTERP_B5, &i_up_b5);
    r_up_b5.gain = 2.0f;
    hipaccLaunchUser(r_up_b5, i_up_b5);
    Resize r_up_cf;
    r_up_cf.in = hipaccDevAccessor<true>(a_cf, HipaccBoundary::CONSTANT, 0.0f, HIPACC_INTERP_CF, &i_up_cf);
    r_up_cf.gain = -1.0f;
    hipaccLaunchUser(r_up_cf, i_up_cf);
    Resize r_up_l3;
    r_up_l3.in = hipaccDevAccessor<true>(a_l3, HipaccBoundary::CLAMP, 0, HIPACC_INTERP_L3, &i_up_l3);
    r_up_l3.gain = 1.5f;
    hipaccLaunchUser(r_up_l3, i_up_l3);

    HipaccAccessor<float> i_dn_nn = hipaccMakeAccessor(dn_nn);
    HipaccAccessor<float> i_dn_lf = hipaccMakeAccessor(dn_lf);
    HipaccAccessor<float> i_dn_b5 = hipaccMakeAccessor(dn_b5);
    HipaccAccessor<float> i_dn_cf = hipaccMakeAccessor(dn_cf);
    HipaccAccessor<float> i_dn_l3 = hipaccMakeAccessor(dn_l3);
    Resize r_dn_nn;
    r_dn_nn.in = hipaccDevAccessor<true>(a_nn, HipaccBoundary::CLAMP, 0, HIPACC_INTERP_NN, &i_dn_nn);
    r_dn_nn.gain = 1.0f;
    hipaccLaunchUser(r_dn_nn, i_dn_nn);
    Resize r_dn_lf;
    r_dn_lf.in = hipaccDevAccessor<true>(a_lf, HipaccBoundary::MIRROR, 0, HIPACC_INTERP_LF, &i_dn_lf);
    r_dn_lf.gain = 0.5f;
    hipaccLaunchUser(r_dn_lf, i_dn_lf);
    Resize r_dn_b5;
    r_dn_b5.in = hipaccDevAccessor<true>(a_b5, HipaccBoundary::REPEAT, 0, HIPACC_INTERP_B5, &i_dn_b5);
    r_dn_b5.gain = 2.0f;
    hipaccLaunchUser(r_dn_b5, i_dn_b5);
    Resize r_dn_cf;
    r_dn_cf.in = hipaccDevAccessor<true>(a_cf, HipaccBoundary::CONSTANT, 0.0f, HIPACC_INTERP_CF, &i_dn_cf);
    r_dn_cf.gain = -1.0f;
    hipaccLaunchUser(r_dn_cf, i_dn_cf);
    Resize r_dn_l3;
    r_dn_l3.in = hipaccDevAccessor<true>(a_l3, HipaccBoundary::CLAMP, 0, HIPACC_INTERP_L3, &i_dn_l3);
    r_dn_l3.gain = 1.5f;
    hipaccLaunchUser(r_dn_l3, i_dn_l3);

    FILE *f = std::fopen(argv[1], "wb");
    std::fwrite(hipaccReadMemory<float>(up_nn), sizeof(float), up_w * up_h, f);
    std::fwrite(hipaccReadMemory<float>(up_lf), sizeof(float), up_w * up_h, f);
    std::fwrite(hipaccReadMemory<float>(up_b5), sizeof(float), up_w * up_h, f);
    std::fwrite(hipaccReadMemory<float>(up_cf), sizeof(float), up_w * up_h, f);
    std::fwrite(hipaccReadMemory<float>(up_l3), sizeof(float), up_w * up_h, f);
    std::fwrite(hipaccReadMemory<float>(dn_nn), sizeof(float), down_w * down_h, f);
    std::fwrite(hipaccReadMemory<float>(dn_lf), sizeof(float), down_w * down_h, f);
    std::fwrite(hipaccReadMemory<float>(dn_b5), sizeof(float), down_w * down_h, f);
    std::fwrite(hipaccReadMemory<float>(dn_cf), sizeof(float), down_w * down_h, f);
    std::fwrite(hipaccReadMemory<float>(dn_l3), sizeof(float), down_w * down_h, f);
    std::fclose(f);
    delete[] pixels;
    return 0;
}
