// The kernel chain of the Mixed fusion sample on uchar images: point, point,
// 5x5 local, point, point, 5x5 local, point, point, each kernel over the
// previous one's image (point: v = in(); v += 3; local: the rounded 5x5
// Gaussian, CLAMP).  The intermediate images are named only by the chain,
// so the emitter launches the chain as one fused kernel.  Writes the last
// image to argv[1].
#include "hipacc.hpp"

#include <cstdint>
#include <cstdio>

using namespace hipacc;

class PointOperator : public Kernel<uchar> {
  private:
    Accessor<uchar> &in;

  public:
    PointOperator(IterationSpace<uchar> &iter, Accessor<uchar> &acc) : Kernel(iter), in(acc) {
        add_accessor(&in);
    }

    void kernel() {
        uchar interm_pixel = in();
        interm_pixel += 3;
        output() = interm_pixel;
    }
};

class LocalOperator : public Kernel<uchar> {
  private:
    Accessor<uchar> &input;
    Mask<float> &mask;

  public:
    LocalOperator(IterationSpace<uchar> &iter, Accessor<uchar> &input, Mask<float> &mask)
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
    uint32_t h = (uint32_t)x * 2654435761u ^ (uint32_t)y * 97u;
    h ^= h >> 15;
    h *= 0x2c1b3c6du;
    return h ^ (h >> 12);
}

HIPACC_CODEGEN int main(int argc, const char **argv) {
    const int width = 523;
    const int height = 317;
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
            input[y * width + x] = (uchar)(pixel_hash(x, y) >> 24);

    Image<uchar> in(width, height, input);
    Image<uchar> out(width, height);
    Mask<float> mask(coef);

    Image<uchar> buf0(width, height);
    Accessor<uchar> acc0(in);
    IterationSpace<uchar> iter0(buf0);
    PointOperator pointOp0(iter0, acc0);

    Image<uchar> buf1(width, height);
    Accessor<uchar> acc1(buf0);
    IterationSpace<uchar> iter1(buf1);
    PointOperator pointOp1(iter1, acc1);

    Image<uchar> buf2(width, height);
    BoundaryCondition<uchar> bound2(buf1, mask, Boundary::CLAMP);
    Accessor<uchar> acc2(bound2);
    IterationSpace<uchar> iter2(buf2);
    LocalOperator localOp2(iter2, acc2, mask);

    Image<uchar> buf3(width, height);
    Accessor<uchar> acc3(buf2);
    IterationSpace<uchar> iter3(buf3);
    PointOperator pointOp3(iter3, acc3);

    Image<uchar> buf4(width, height);
    Accessor<uchar> acc4(buf3);
    IterationSpace<uchar> iter4(buf4);
    PointOperator pointOp4(iter4, acc4);

    Image<uchar> buf5(width, height);
    BoundaryCondition<uchar> bound5(buf4, mask, Boundary::CLAMP);
    Accessor<uchar> acc5(bound5);
    IterationSpace<uchar> iter5(buf5);
    LocalOperator localOp5(iter5, acc5, mask);

    Image<uchar> buf6(width, height);
    Accessor<uchar> acc6(buf5);
    IterationSpace<uchar> iter6(buf6);
    PointOperator pointOp6(iter6, acc6);

    Accessor<uchar> acc7(buf6);
    IterationSpace<uchar> iter7(out);
    PointOperator pointOp7(iter7, acc7);

    pointOp0.execute();
    pointOp1.execute();
    localOp2.execute();
    pointOp3.execute();
    pointOp4.execute();
    localOp5.execute();
    pointOp6.execute();
    pointOp7.execute();

    FILE *f = std::fopen(argv[1], "wb");
    std::fwrite(out.data(), 1, (size_t)width * height, f);
    std::fclose(f);
    delete[] input;
    return 0;
}
