// Global reductions of a float image (Kernel::reduce and reduced_data(), in
// the form of the Reduction samples): minimum, maximum and sum.  Writes the
// three results to argv[1], then the sum and the sum of magnitudes in double
// and the pixel count.
#include "hipacc.hpp"

#include <cstdint>
#include <cstdio>

using namespace hipacc;
using namespace hipacc::math;

class MinReduction : public Kernel<float> {
  private:
    Accessor<float> &input;

  public:
    MinReduction(IterationSpace<float> &iter, Accessor<float> &input) : Kernel(iter), input(input) {
        add_accessor(&input);
    }

    void kernel() { output() = input(); }

    float reduce(float left, float right) const { return min(left, right); }
};

class MaxReduction : public Kernel<float> {
  private:
    Accessor<float> &input;

  public:
    MaxReduction(IterationSpace<float> &iter, Accessor<float> &input) : Kernel(iter), input(input) {
        add_accessor(&input);
    }

    void kernel() { output() = input(); }

    float reduce(float left, float right) const { return max(left, right); }
};

class SumReduction : public Kernel<float> {
  private:
    Accessor<float> &input;

  public:
    SumReduction(IterationSpace<float> &iter, Accessor<float> &input) : Kernel(iter), input(input) {
        add_accessor(&input);
    }

    void kernel() { output() = input(); }

    float reduce(float left, float right) const { return left + right; }
};

static float pixel_value(int x, int y) {
    uint32_t h = (uint32_t)x * 2246822519u + (uint32_t)y * 3266489917u;
    h = (h ^ (h >> 15)) * 668265263u;
    return (float)((h ^ (h >> 13)) >> 8) * (1.0f / 16777216.0f) - 0.25f;
}

HIPACC_CODEGEN int main(int argc, const char **argv) {
    const int width = 1024;
    const int height = 333;
    float *input = new float[width * height];
    for (int y = 0; y < height; ++y)
        for (int x = 0; x < width; ++x)
            input[y * width + x] = pixel_value(x, y);

    Image<float> in(width, height, input);
    Image<float> out(width, height);
    Accessor<float> acc(in);
    IterationSpace<float> iter(out);

    MinReduction rmin(iter, acc);
    rmin.execute();
    float vmin = rmin.reduced_data();
    MaxReduction rmax(iter, acc);
    rmax.execute();
    float vmax = rmax.reduced_data();
    SumReduction rsum(iter, acc);
    rsum.execute();
    float vsum = rsum.reduced_data();

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
