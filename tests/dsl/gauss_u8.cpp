// Gaussian blur of a uchar image: 5x5 mask, CLAMP, (uchar)(sum + 0.5f), in
// the form of the Gaussian_Blur sample.  Writes the output image to argv[1].
#include "hipacc.hpp"

#include <cstdint>
#include <cstdio>

using namespace hipacc;

class GaussianBlur : public Kernel<uchar> {
  private:
    Accessor<uchar> &input;
    Mask<float> &mask;

  public:
    GaussianBlur(IterationSpace<uchar> &iter, Accessor<uchar> &input, Mask<float> &mask)
        : Kernel(iter), input(input), mask(mask) {
        add_accessor(&input);
    }

    void kernel() {
        output() = (uchar)(convolve(mask, Reduce::SUM, [&] () -> float {
                return mask() * input(mask);
            }) + 0.5f);
    }
};

static uint32_t pixel_hash(int x, int y) {
    uint32_t h = (uint32_t)x * 73856093u ^ (uint32_t)y * 19349663u;
    h ^= h >> 13;
    h *= 0x5bd1e995u;
    return h ^ (h >> 15);
}

HIPACC_CODEGEN int main(int argc, const char **argv) {
    const int width = 1000;
    const int height = 600;
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
            input[y * width + x] = (uchar)(pixel_hash(x, y) & 255);

    Image<uchar> in(width, height, input);
    Image<uchar> out(width, height);
    Mask<float> mask(coef);
    BoundaryCondition<uchar> bound(in, mask, Boundary::CLAMP);
    Accessor<uchar> acc(bound);
    IterationSpace<uchar> iter(out);
    GaussianBlur filter(iter, acc, mask);
    filter.execute();
    uchar *output = out.data();

    FILE *f = std::fopen(argv[1], "wb");
    std::fwrite(output, 1, (size_t)width * height, f);
    std::fclose(f);
    delete[] input;
    return 0;
}
