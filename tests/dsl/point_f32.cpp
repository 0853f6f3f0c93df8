// Point operators and a float local operator on float images: a scale and
// offset, the fusion samples' `v += 3` point kernel, a difference of two
// images and a 3x3 float convolution (MIRROR) of the difference.  Writes the
// outputs to argv[1], one after the other.
#include "hipacc.hpp"

#include <cstdint>
#include <cstdio>

using namespace hipacc;

class Scale : public Kernel<float> {
  private:
    Accessor<float> &input;

  public:
    Scale(IterationSpace<float> &iter, Accessor<float> &input) : Kernel(iter), input(input) {
        add_accessor(&input);
    }

    void kernel() { output() = input() * 0.5f + 0.25f; }
};

class AddThree : public Kernel<float> {
  private:
    Accessor<float> &input;

  public:
    AddThree(IterationSpace<float> &iter, Accessor<float> &input) : Kernel(iter), input(input) {
        add_accessor(&input);
    }

    void kernel() {
        float v = input();
        v += 3;
        output() = v;
    }
};

class Difference : public Kernel<float> {
  private:
    Accessor<float> &a;
    Accessor<float> &b;

  public:
    Difference(IterationSpace<float> &iter, Accessor<float> &a, Accessor<float> &b)
        : Kernel(iter), a(a), b(b) {
        add_accessor(&a);
        add_accessor(&b);
    }

    void kernel() { output() = a() - b(); }
};

class Blur3 : public Kernel<float> {
  private:
    Accessor<float> &input;
    Mask<float> &mask;

  public:
    Blur3(IterationSpace<float> &iter, Accessor<float> &input, Mask<float> &mask)
        : Kernel(iter), input(input), mask(mask) {
        add_accessor(&input);
    }

    void kernel() {
        output() = convolve(mask, Reduce::SUM, [&] () -> float { return mask() * input(mask); });
    }
};

static float pixel_value(int x, int y) {
    uint32_t h = (uint32_t)x * 374761393u + (uint32_t)y * 668265263u;
    h = (h ^ (h >> 13)) * 1274126177u;
    return (float)((h ^ (h >> 16)) >> 8) * (1.0f / 16777216.0f);
}

HIPACC_CODEGEN int main(int argc, const char **argv) {
    const int width = 640;
    const int height = 257;
    const float coef[3][3] = { { 0.0625f, 0.125f, 0.0625f },
                               { 0.125f, 0.25f, 0.125f },
                               { 0.0625f, 0.125f, 0.0625f } };
    float *input = new float[width * height];
    for (int y = 0; y < height; ++y)
        for (int x = 0; x < width; ++x)
            input[y * width + x] = pixel_value(x, y);

    Image<float> in(width, height, input);
    Image<float> scaled(width, height);
    Image<float> shifted(width, height);
    Image<float> diff(width, height);
    Image<float> blurred(width, height);

    Accessor<float> ain(in);
    IterationSpace<float> iscaled(scaled);
    Scale scale(iscaled, ain);
    scale.execute();

    Accessor<float> ascaled(scaled);
    IterationSpace<float> ishifted(shifted);
    AddThree add_three(ishifted, ascaled);
    add_three.execute();

    Accessor<float> ashifted(shifted);
    IterationSpace<float> idiff(diff);
    Difference difference(idiff, ashifted, ain);
    difference.execute();

    Mask<float> mask(coef);
    BoundaryCondition<float> bdiff(diff, mask, Boundary::MIRROR);
    Accessor<float> adiff(bdiff);
    IterationSpace<float> iblurred(blurred);
    Blur3 blur(iblurred, adiff, mask);
    blur.execute();

    FILE *f = std::fopen(argv[1], "wb");
    std::fwrite(scaled.data(), sizeof(float), (size_t)width * height, f);
    std::fwrite(shifted.data(), sizeof(float), (size_t)width * height, f);
    std::fwrite(diff.data(), sizeof(float), (size_t)width * height, f);
    std::fwrite(blurred.data(), sizeof(float), (size_t)width * height, f);
    std::fclose(f);
    delete[] input;
    return 0;
}
