// Kernels the library has no operator for, in the form of the Unsharp and
// Laplace samples: a per-pixel polynomial, a weighted difference of two
// images, a ratio, and a 5-point Laplacian written with explicit offsets
// (MIRROR); the 5x5 blur between them is a library convolution.  Writes the
// outputs to argv[1], one after the other.
#include "hipacc.hpp"

#include <cstdint>
#include <cstdio>

using namespace hipacc;

class Tone : public Kernel<float> {
  private:
    Accessor<float> &input;

  public:
    Tone(IterationSpace<float> &iter, Accessor<float> &input) : Kernel(iter), input(input) {
        add_accessor(&input);
    }

    void kernel() {
        float v = input();
        output() = v * v * 0.5f + v * 0.5f;
    }
};

class Blur : public Kernel<float> {
  private:
    Accessor<float> &input;
    Mask<float> &mask;

  public:
    Blur(IterationSpace<float> &iter, Accessor<float> &input, Mask<float> &mask)
        : Kernel(iter), input(input), mask(mask) {
        add_accessor(&input);
    }

    void kernel() {
        output() = convolve(mask, Reduce::SUM, [&] () -> float { return mask() * input(mask); });
    }
};

class Sharpen : public Kernel<float> {
  private:
    Accessor<float> &Gray;
    Accessor<float> &InputBlur;

  public:
    Sharpen(IterationSpace<float> &iter, Accessor<float> &Gray, Accessor<float> &InputBlur)
        : Kernel(iter), Gray(Gray), InputBlur(InputBlur) {
        add_accessor(&Gray);
        add_accessor(&InputBlur);
    }

    void kernel() { output() = 2 * Gray() - InputBlur(); }
};

class Ratio : public Kernel<float> {
  private:
    Accessor<float> &Sharp;
    Accessor<float> &Gray;
    float eps;

  public:
    Ratio(IterationSpace<float> &iter, Accessor<float> &Sharp, Accessor<float> &Gray, float eps)
        : Kernel(iter), Sharp(Sharp), Gray(Gray), eps(eps) {
        add_accessor(&Sharp);
        add_accessor(&Gray);
    }

    void kernel() { output() = Sharp() / (Gray() + eps); }
};

class Laplace : public Kernel<float> {
  private:
    Accessor<float> &input;

  public:
    Laplace(IterationSpace<float> &iter, Accessor<float> &input) : Kernel(iter), input(input) {
        add_accessor(&input);
    }

    void kernel() {
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

    Image<float> in(width, height, input);
    Image<float> gray(width, height);
    Image<float> blurred(width, height);
    Image<float> sharp(width, height);
    Image<float> ratio(width, height);
    Image<float> lap(width, height);

    Accessor<float> ain(in);
    IterationSpace<float> igray(gray);
    Tone tone(igray, ain);
    tone.execute();

    Mask<float> mask(coef);
    BoundaryCondition<float> bgray(gray, mask, Boundary::CLAMP);
    Accessor<float> agray_bc(bgray);
    IterationSpace<float> iblurred(blurred);
    Blur blur(iblurred, agray_bc, mask);
    blur.execute();

    Accessor<float> agray(gray);
    Accessor<float> ablurred(blurred);
    IterationSpace<float> isharp(sharp);
    Sharpen sharpen(isharp, agray, ablurred);
    sharpen.execute();

    Accessor<float> asharp(sharp);
    IterationSpace<float> iratio(ratio);
    Ratio rat(iratio, asharp, agray, 0.01f);
    rat.execute();

    BoundaryCondition<float> bmirror(in, 3, 3, Boundary::MIRROR);
    Accessor<float> amirror(bmirror);
    IterationSpace<float> ilap(lap);
    Laplace laplace(ilap, amirror);
    laplace.execute();

    FILE *f = std::fopen(argv[1], "wb");
    std::fwrite(gray.data(), sizeof(float), (size_t)width * height, f);
    std::fwrite(blurred.data(), sizeof(float), (size_t)width * height, f);
    std::fwrite(sharp.data(), sizeof(float), (size_t)width * height, f);
    std::fwrite(ratio.data(), sizeof(float), (size_t)width * height, f);
    std::fwrite(lap.data(), sizeof(float), (size_t)width * height, f);
    std::fclose(f);
    delete[] input;
    return 0;
}
