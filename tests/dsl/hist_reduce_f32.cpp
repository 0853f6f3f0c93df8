// Global operators the library has no fixed operator for, in the form of the
// Histogram sample (Kernel<float, uint> with binning() and reduce()) and a
// reduction whose reduce() keeps the value of largest magnitude.  Writes the
// 64 bins and the reduction's value to argv[1].
#include "hipacc.hpp"

#include <cmath>
#include <cstdint>
#include <cstdio>

using namespace hipacc;
using namespace hipacc::math;

class Histogram : public Kernel<float, uint> {
  private:
    Accessor<float> &in;

  public:
    Histogram(IterationSpace<float> &iter, Accessor<float> &in) : Kernel(iter), in(in) {
        add_accessor(&in);
    }

    void kernel() { output() = in(); }

    void binning(unsigned int x, unsigned int y, float pixel) {
        bin(pixel / 255.0f * num_bins()) = 1;
    }

    uint reduce(uint left, uint right) const { return left + right; }
};

class MaxMagnitude : public Kernel<float> {
  private:
    Accessor<float> &in;

  public:
    MaxMagnitude(IterationSpace<float> &iter, Accessor<float> &in) : Kernel(iter), in(in) {
        add_accessor(&in);
    }

    void kernel() { output() = in() - 100.0f; }

    float reduce(float left, float right) const {
        return fabsf(left) > fabsf(right) ? left : right;
    }
};

static float pixel_value(int x, int y) {
    uint32_t h = (uint32_t)x * 3266489917u + (uint32_t)y * 668265263u;
    h = (h ^ (h >> 15)) * 2246822519u;
    return (float)((h ^ (h >> 13)) >> 8) * (255.0f / 16777216.0f);
}

HIPACC_CODEGEN int main(int argc, const char **argv) {
    const int width = 517;
    const int height = 263;
    const int num_bins = 64;
    float *input = new float[width * height];
    for (int y = 0; y < height; ++y)
        for (int x = 0; x < width; ++x)
            input[y * width + x] = pixel_value(x, y);

    Image<float> in(width, height, input);
    Image<float> copy(width, height);
    Image<float> shifted(width, height);

    Accessor<float> acc(in);
    IterationSpace<float> icopy(copy);
    Histogram hist(icopy, acc);
    hist.execute();
    uint *bins = hist.binned_data(num_bins);

    IterationSpace<float> ishifted(shifted);
    MaxMagnitude maxmag(ishifted, acc);
    maxmag.execute();
    float extreme = maxmag.reduced_data();

    FILE *f = std::fopen(argv[1], "wb");
    std::fwrite(bins, sizeof(uint), num_bins, f);
    std::fwrite(&extreme, sizeof(float), 1, f);
    std::fclose(f);
    delete[] bins;
    delete[] input;
    return 0;
}
