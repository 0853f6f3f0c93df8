// Interpolating accessors read by kernels that are not library operators:
// a float image scaled up (x 1.7, x 1.3) and down (x 0.6) through accessors
// with Interpolate::NN, LF, B5, CF and L3, each over a boundary condition
// (CLAMP, MIRROR, REPEAT, CONSTANT), read at the pixel and at an offset,
// with the result scaled and offset in the kernel.  Writes the up- and
// down-scaled images of every interpolation mode to argv[1].
#include "hipacc.hpp"

#include <cstdint>
#include <cstdio>

using namespace hipacc;
using namespace hipacc::math;

class Resize : public Kernel<float> {
  private:
    Accessor<float> &in;
    float gain;

  public:
    Resize(IterationSpace<float> &iter, Accessor<float> &in, float gain)
        : Kernel(iter), in(in), gain(gain) {
        add_accessor(&in);
    }

    void kernel() { output() = gain * in() + 0.25f * in(1, -1); }
};

static float value(int x, int y) {
    uint32_t h = (uint32_t)x * 2654435761u ^ (uint32_t)y * 2246822519u;
    h = (h ^ (h >> 16)) * 2146121005u;
    return (float)((h ^ (h >> 15)) >> 8) * (1.0f / 16777216.0f) - 0.5f;
}

HIPACC_CODEGEN int main(int argc, const char **argv) {
    const int width = 123;
    const int height = 77;
    const int up_w = 209, up_h = 100;     // x 1.7, x 1.3
    const int down_w = 74, down_h = 46;   // x 0.6
    const float dummy[3][3] = { { 0.f, 0.f, 0.f }, { 0.f, 1.f, 0.f }, { 0.f, 0.f, 0.f } };
    float *pixels = new float[width * height];
    for (int y = 0; y < height; ++y)
        for (int x = 0; x < width; ++x)
            pixels[y * width + x] = value(x, y);

    Image<float> in(width, height, pixels);
    Mask<float> mask(dummy);
    Image<float> up_nn(up_w, up_h), up_lf(up_w, up_h), up_b5(up_w, up_h), up_cf(up_w, up_h),
        up_l3(up_w, up_h);
    Image<float> dn_nn(down_w, down_h), dn_lf(down_w, down_h), dn_b5(down_w, down_h),
        dn_cf(down_w, down_h), dn_l3(down_w, down_h);

    BoundaryCondition<float> b_clamp(in, mask, Boundary::CLAMP);
    BoundaryCondition<float> b_mirror(in, mask, Boundary::MIRROR);
    BoundaryCondition<float> b_repeat(in, mask, Boundary::REPEAT);
    // the DSL host emulation reads 0 whatever constant is given (its
    // BoundaryCondition never stores the value, dsl/image.hpp:259-290), so
    // the constant here is 0
    BoundaryCondition<float> b_const(in, mask, Boundary::CONSTANT, 0.0f);

    Accessor<float> a_nn(b_clamp, Interpolate::NN);
    Accessor<float> a_lf(b_mirror, Interpolate::LF);
    Accessor<float> a_b5(b_repeat, Interpolate::B5);
    Accessor<float> a_cf(b_const, Interpolate::CF);
    Accessor<float> a_l3(b_clamp, Interpolate::L3);

    IterationSpace<float> i_up_nn(up_nn), i_up_lf(up_lf), i_up_b5(up_b5), i_up_cf(up_cf),
        i_up_l3(up_l3);
    Resize r_up_nn(i_up_nn, a_nn, 1.0f);
    r_up_nn.execute();
    Resize r_up_lf(i_up_lf, a_lf, 0.5f);
    r_up_lf.execute();
    Resize r_up_b5(i_up_b5, a_b5, 2.0f);
    r_up_b5.execute();
    Resize r_up_cf(i_up_cf, a_cf, -1.0f);
    r_up_cf.execute();
    Resize r_up_l3(i_up_l3, a_l3, 1.5f);
    r_up_l3.execute();

    IterationSpace<float> i_dn_nn(dn_nn), i_dn_lf(dn_lf), i_dn_b5(dn_b5), i_dn_cf(dn_cf),
        i_dn_l3(dn_l3);
    Resize r_dn_nn(i_dn_nn, a_nn, 1.0f);
    r_dn_nn.execute();
    Resize r_dn_lf(i_dn_lf, a_lf, 0.5f);
    r_dn_lf.execute();
    Resize r_dn_b5(i_dn_b5, a_b5, 2.0f);
    r_dn_b5.execute();
    Resize r_dn_cf(i_dn_cf, a_cf, -1.0f);
    r_dn_cf.execute();
    Resize r_dn_l3(i_dn_l3, a_l3, 1.5f);
    r_dn_l3.execute();

    FILE *f = std::fopen(argv[1], "wb");
    std::fwrite(up_nn.data(), sizeof(float), up_w * up_h, f);
    std::fwrite(up_lf.data(), sizeof(float), up_w * up_h, f);
    std::fwrite(up_b5.data(), sizeof(float), up_w * up_h, f);
    std::fwrite(up_cf.data(), sizeof(float), up_w * up_h, f);
    std::fwrite(up_l3.data(), sizeof(float), up_w * up_h, f);
    std::fwrite(dn_nn.data(), sizeof(float), down_w * down_h, f);
    std::fwrite(dn_lf.data(), sizeof(float), down_w * down_h, f);
    std::fwrite(dn_b5.data(), sizeof(float), down_w * down_h, f);
    std::fwrite(dn_cf.data(), sizeof(float), down_w * down_h, f);
    std::fwrite(dn_l3.data(), sizeof(float), down_w * down_h, f);
    std::fclose(f);
    delete[] pixels;
    return 0;
}
