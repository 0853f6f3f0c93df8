// Integer local operators on one uchar image: Sobel-x with an int mask over
// Domain(mask) into short (MIRROR), a 3x3 erode (REPEAT) and a 5x5 box sum
// into uint (CONSTANT).  Writes the three outputs to argv[1], one after the
// other.
#include "hipacc.hpp"

#include <cstdint>
#include <cstdio>

using namespace hipacc;

class SobelX : public Kernel<short> {
  private:
    Accessor<uchar> &input;
    Domain &dom;
    Mask<int> &mask;

  public:
    SobelX(IterationSpace<short> &iter, Accessor<uchar> &input, Domain &dom, Mask<int> &mask)
        : Kernel(iter), input(input), dom(dom), mask(mask) {
        add_accessor(&input);
    }

    void kernel() {
        output() = (short)reduce(dom, Reduce::SUM, [&] () -> int {
            return mask(dom) * input(dom);
        });
    }
};

class Erode : public Kernel<uchar> {
  private:
    Accessor<uchar> &input;
    Domain &dom;

  public:
    Erode(IterationSpace<uchar> &iter, Accessor<uchar> &input, Domain &dom)
        : Kernel(iter), input(input), dom(dom) {
        add_accessor(&input);
    }

    void kernel() {
        output() = reduce(dom, Reduce::MIN, [&] () -> uchar { return input(dom); });
    }
};

class BoxSum : public Kernel<uint> {
  private:
    Accessor<uchar> &input;
    Domain &dom;

  public:
    BoxSum(IterationSpace<uint> &iter, Accessor<uchar> &input, Domain &dom)
        : Kernel(iter), input(input), dom(dom) {
        add_accessor(&input);
    }

    void kernel() {
        output() = reduce(dom, Reduce::SUM, [&] () -> int { return input(dom); });
    }
};

static uint32_t pixel_hash(int x, int y) {
    uint32_t h = (uint32_t)x * 2654435761u + (uint32_t)y * 40503u;
    h ^= h >> 16;
    h *= 0x7feb352du;
    return h ^ (h >> 15);
}

HIPACC_CODEGEN int main(int argc, const char **argv) {
    const int width = 777;
    const int height = 301;
    const int sobel[3][3] = { { -1, 0, 1 }, { -2, 0, 2 }, { -1, 0, 1 } };
    uchar *input = new uchar[width * height];
    for (int y = 0; y < height; ++y)
        for (int x = 0; x < width; ++x)
            input[y * width + x] = (uchar)(pixel_hash(x, y) >> 24);

    Image<uchar> in(width, height, input);
    Image<short> grad(width, height);
    Image<uchar> eroded(width, height);
    Image<uint> box(width, height);

    Mask<int> smask(sobel);
    Domain sdom(smask);
    BoundaryCondition<uchar> bmirror(in, smask, Boundary::MIRROR);
    Accessor<uchar> amirror(bmirror);
    IterationSpace<short> igrad(grad);
    SobelX sobel_x(igrad, amirror, sdom, smask);
    sobel_x.execute();

    Domain d3(3, 3);
    BoundaryCondition<uchar> brepeat(in, d3, Boundary::REPEAT);
    Accessor<uchar> arepeat(brepeat);
    IterationSpace<uchar> ieroded(eroded);
    Erode erode(ieroded, arepeat, d3);
    erode.execute();

    Domain d5(5, 5);
    BoundaryCondition<uchar> bconst(in, d5, Boundary::CONSTANT, 0);
    Accessor<uchar> aconst(bconst);
    IterationSpace<uint> ibox(box);
    BoxSum box_sum(ibox, aconst, d5);
    box_sum.execute();

    FILE *f = std::fopen(argv[1], "wb");
    std::fwrite(grad.data(), sizeof(short), (size_t)width * height, f);
    std::fwrite(eroded.data(), 1, (size_t)width * height, f);
    std::fwrite(box.data(), sizeof(uint), (size_t)width * height, f);
    std::fclose(f);
    delete[] input;
    return 0;
}
