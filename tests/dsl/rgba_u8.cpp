// Kernels on 4-wide uchar4 pixels, in the form of the RGBA samples and of
// the motion-interpolation sample's vector binning: a float4 weighted sum
// through convert_float4 / convert_uchar4 (MIRROR), a componentwise MIN
// over a domain (CLAMP), an int4 contrast stretch with vector-scalar
// arithmetic and min / max, and per-tile float4 bins (sums of the channels
// and a count) combined by a reduce() of float4 values.  Every bin sum is an
// integer below 2^24, so it is exact in any order.  Writes the three images
// and the bins to argv[1].
#include "hipacc.hpp"

#include <cstdint>
#include <cstdio>

using namespace hipacc;
using namespace hipacc::math;

class Smooth : public Kernel<uchar4> {
  private:
    Accessor<uchar4> &in;
    Mask<float> &mask;

  public:
    Smooth(IterationSpace<uchar4> &iter, Accessor<uchar4> &in, Mask<float> &mask)
        : Kernel(iter), in(in), mask(mask) {
        add_accessor(&in);
    }

    void kernel() {
        float4 s = convolve(mask, Reduce::SUM, [&] () -> float4 {
            return mask() * convert_float4(in(mask));
        });
        output() = convert_uchar4(s + 0.5f);
    }
};

class Shrink : public Kernel<uchar4> {
  private:
    Accessor<uchar4> &in;
    Domain &dom;

  public:
    Shrink(IterationSpace<uchar4> &iter, Accessor<uchar4> &in, Domain &dom)
        : Kernel(iter), in(in), dom(dom) {
        add_accessor(&in);
    }

    void kernel() {
        output() = reduce(dom, Reduce::MIN, [&] () -> uchar4 { return in(dom); });
    }
};

class Stretch : public Kernel<uchar4> {
  private:
    Accessor<uchar4> &in;

  public:
    Stretch(IterationSpace<uchar4> &iter, Accessor<uchar4> &in) : Kernel(iter), in(in) {
        add_accessor(&in);
    }

    void kernel() {
        int4 d = convert_int4(in()) - 128;
        output() = convert_uchar4(max(min(d * 3 + 128, 255), 0));
    }
};

class TileSums : public Kernel<uchar4, float4> {
  private:
    Accessor<uchar4> &in;
    int tiles_x;

  public:
    TileSums(IterationSpace<uchar4> &iter, Accessor<uchar4> &in, int tiles_x)
        : Kernel(iter), in(in), tiles_x(tiles_x) {
        add_accessor(&in);
    }

    void kernel() { output() = in(); }

    void binning(unsigned int x, unsigned int y, uchar4 pixel) {
        float4 v = convert_float4(pixel);
        v.w = 1.0f;
        bin((y / 16) * tiles_x + x / 16) = v;
    }

    float4 reduce(float4 left, float4 right) const { return left + right; }
};

static uint32_t mix(int x, int y, uint32_t salt) {
    uint32_t h = (uint32_t)x * 2654435761u ^ (uint32_t)y * 2246822519u ^ salt;
    h = (h ^ (h >> 16)) * 2146121005u;
    return h ^ (h >> 15);
}

HIPACC_CODEGEN int main(int argc, const char **argv) {
    const int width = 197;
    const int height = 131;
    const int tiles_x = (width + 15) / 16;
    const int tiles_y = (height + 15) / 16;
    const float coef[3][5] = {
        { 0.02f, 0.05f, 0.08f, 0.05f, 0.02f },
        { 0.05f, 0.10f, 0.16f, 0.10f, 0.05f },
        { 0.02f, 0.05f, 0.08f, 0.05f, 0.02f }
    };
    const uchar cross[3][3] = {
        { 0, 1, 0 },
        { 1, 1, 1 },
        { 0, 1, 0 }
    };
    uchar4 *pixels = new uchar4[width * height];
    for (int y = 0; y < height; ++y)
        for (int x = 0; x < width; ++x) {
            uint32_t h = mix(x, y, 3u);
            pixels[y * width + x] = make_uchar4((uchar)h, (uchar)(h >> 8), (uchar)(h >> 16),
                                                (uchar)(h >> 24));
        }

    Image<uchar4> in(width, height, pixels);
    Image<uchar4> smooth(width, height);
    Image<uchar4> shrink(width, height);
    Image<uchar4> stretch(width, height);
    Image<uchar4> copy(width, height);
    Mask<float> mask(coef);
    Domain dom(cross);

    BoundaryCondition<uchar4> bc_mirror(in, mask, Boundary::MIRROR);
    Accessor<uchar4> a_mirror(bc_mirror);
    IterationSpace<uchar4> i_smooth(smooth);
    Smooth k_smooth(i_smooth, a_mirror, mask);
    k_smooth.execute();

    BoundaryCondition<uchar4> bc_clamp(in, dom, Boundary::CLAMP);
    Accessor<uchar4> a_clamp(bc_clamp);
    IterationSpace<uchar4> i_shrink(shrink);
    Shrink k_shrink(i_shrink, a_clamp, dom);
    k_shrink.execute();

    Accessor<uchar4> a_in(in);
    IterationSpace<uchar4> i_stretch(stretch);
    Stretch k_stretch(i_stretch, a_in);
    k_stretch.execute();

    IterationSpace<uchar4> i_copy(copy);
    TileSums k_tiles(i_copy, a_in, tiles_x);
    float4 *bins = k_tiles.binned_data(tiles_x * tiles_y);

    FILE *f = std::fopen(argv[1], "wb");
    std::fwrite(smooth.data(), sizeof(uchar4), width * height, f);
    std::fwrite(shrink.data(), sizeof(uchar4), width * height, f);
    std::fwrite(stretch.data(), sizeof(uchar4), width * height, f);
    std::fwrite(bins, sizeof(float4), tiles_x * tiles_y, f);
    std::fclose(f);
    delete[] bins;
    delete[] pixels;
    return 0;
}
