// The bilateral filter in the form of the Bilateral_Filter sample: a
// closeness mask computed on the host, a Domain over it, and iterate() with
// the range weight expf(-c_r * diff^2) per tap (CLAMP).  Writes the output
// image to argv[1].
#include "hipacc.hpp"

#include <cmath>
#include <cstdint>
#include <cstdio>

using namespace hipacc;

class BilateralFilter : public Kernel<uchar> {
  private:
    Accessor<uchar> &in;
    Mask<float> &mask;
    Domain &dom;
    int sigma_r;

  public:
    BilateralFilter(IterationSpace<uchar> &iter, Accessor<uchar> &in, Mask<float> &mask,
                    Domain &dom, int sigma_r)
        : Kernel(iter), in(in), mask(mask), dom(dom), sigma_r(sigma_r) {
        add_accessor(&in);
    }

    void kernel() {
        float c_r = 0.5f / (sigma_r * sigma_r);
        float d = 0.0f;
        float p = 0.0f;
        float center = in();

        iterate(dom, [&] () -> void {
            float diff = (float)(in(dom)) - center;
            float s = expf(-c_r * diff * diff) * mask(dom);
            d += s;
            p += s * in(dom);
        });

        output() = (uchar)(p / d + 0.5f);
    }
};

static uint32_t pixel_hash(int x, int y) {
    uint32_t h = (uint32_t)(x / 3) * 73856093u ^ (uint32_t)(y / 3) * 19349663u;
    h ^= h >> 13;
    h *= 0x5bd1e995u;
    return (h ^ (h >> 15)) + (uint32_t)(x + y);
}

HIPACC_CODEGEN int main(int argc, const char **argv) {
    const int width = 301;
    const int height = 199;
    const int sigma_d = 2;
    const int sigma_r = 16;
    float coef[7][7];
    for (int yf = -3; yf <= 3; ++yf)
        for (int xf = -3; xf <= 3; ++xf)
            coef[yf + 3][xf + 3] = expf(-0.5f / (sigma_d * sigma_d) * (xf * xf + yf * yf));
    uchar *input = new uchar[width * height];
    for (int y = 0; y < height; ++y)
        for (int x = 0; x < width; ++x)
            input[y * width + x] = (uchar)(pixel_hash(x, y) & 255);

    Image<uchar> in(width, height, input);
    Image<uchar> out(width, height);
    Mask<float> mask(coef);
    Domain dom(mask);
    BoundaryCondition<uchar> bound(in, mask, Boundary::CLAMP);
    Accessor<uchar> acc(bound);
    IterationSpace<uchar> iter(out);
    BilateralFilter filter(iter, acc, mask, dom, sigma_r);
    filter.execute();

    FILE *f = std::fopen(argv[1], "wb");
    std::fwrite(out.data(), 1, (size_t)width * height, f);
    std::fclose(f);
    delete[] input;
    return 0;
}
