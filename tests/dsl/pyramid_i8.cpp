// Multi-resolution kernels on char images of an odd size, in the form of the
// Gaussian / Laplacian pyramid sample: a recursive traverse() over three
// pyramids with a convolution (a library operator), nearest-neighbour
// subsampling and linear-filtered upsampling through interpolating
// accessors (device code); then a loop that runs a kernel three times (an
// instance declared in the loop) with an Image = Image copy between the
// runs, and a MAX over a 3x3 domain whose centre is switched off with
// dom(0, 0) = 0.  Level 0 of the pyramids is copied into images of its own
// inside the traversal, so the outputs do not depend on whether a pyramid
// shares its base image.  Writes the Laplacian level 0, the restored image,
// the loop's result and the MAX image to argv[1].
#include "hipacc.hpp"

#include <cstdint>
#include <cstdio>

using namespace hipacc;
using namespace hipacc::math;

class Blur : public Kernel<char> {
  private:
    Accessor<char> &in;
    Mask<float> &mask;

  public:
    Blur(IterationSpace<char> &iter, Accessor<char> &in, Mask<float> &mask)
        : Kernel(iter), in(in), mask(mask) {
        add_accessor(&in);
    }

    void kernel() {
        output() = convolve(mask, Reduce::SUM, [&] () -> float { return mask() * in(mask); });
    }
};

class Take : public Kernel<char> {
  private:
    Accessor<char> &in;

  public:
    Take(IterationSpace<char> &iter, Accessor<char> &in) : Kernel(iter), in(in) {
        add_accessor(&in);
    }

    void kernel() { output() = in(); }
};

class Detail : public Kernel<char> {
  private:
    Accessor<char> &fine;
    Accessor<char> &coarse;

  public:
    Detail(IterationSpace<char> &iter, Accessor<char> &fine, Accessor<char> &coarse)
        : Kernel(iter), fine(fine), coarse(coarse) {
        add_accessor(&fine);
        add_accessor(&coarse);
    }

    void kernel() { output() = fine() - coarse(); }
};

class Rebuild : public Kernel<char> {
  private:
    Accessor<char> &coarse;
    Accessor<char> &detail;

  public:
    Rebuild(IterationSpace<char> &iter, Accessor<char> &coarse, Accessor<char> &detail)
        : Kernel(iter), coarse(coarse), detail(detail) {
        add_accessor(&coarse);
        add_accessor(&detail);
    }

    void kernel() { output() = coarse() + detail(); }
};

class Step : public Kernel<short> {
  private:
    Accessor<short> &in;

  public:
    Step(IterationSpace<short> &iter, Accessor<short> &in) : Kernel(iter), in(in) {
        add_accessor(&in);
    }

    void kernel() {
        short v = in();
        output() = v > 0 ? v / 2 : v + 11;
    }
};

class RingMax : public Kernel<char> {
  private:
    Accessor<char> &in;
    Domain &dom;

  public:
    RingMax(IterationSpace<char> &iter, Accessor<char> &in, Domain &dom)
        : Kernel(iter), in(in), dom(dom) {
        add_accessor(&in);
    }

    void kernel() {
        output() = reduce(dom, Reduce::MAX, [&] () -> char { return in(dom); });
    }
};

static uint32_t mix(int x, int y, uint32_t salt) {
    uint32_t h = (uint32_t)x * 2654435761u ^ (uint32_t)y * 2246822519u ^ salt;
    h = (h ^ (h >> 16)) * 2146121005u;
    return h ^ (h >> 15);
}

HIPACC_CODEGEN int main(int argc, const char **argv) {
    const int width = 301;
    const int height = 197;
    const int depth = 4;
    const float coef[5][5] = {
        { 0.0030f, 0.0133f, 0.0219f, 0.0133f, 0.0030f },
        { 0.0133f, 0.0596f, 0.0983f, 0.0596f, 0.0133f },
        { 0.0219f, 0.0983f, 0.1621f, 0.0983f, 0.0219f },
        { 0.0133f, 0.0596f, 0.0983f, 0.0596f, 0.0133f },
        { 0.0030f, 0.0133f, 0.0219f, 0.0133f, 0.0030f }
    };
    char *pixels = new char[width * height];
    short *start = new short[width * height];
    for (int y = 0; y < height; ++y)
        for (int x = 0; x < width; ++x) {
            pixels[y * width + x] = (char)(mix(x, y, 7u) >> 24);
            start[y * width + x] = (short)((int)(mix(x, y, 11u) >> 20) - 2048);
        }

    Image<char> gauss(width, height, pixels);
    Image<char> blurred(width, height);
    Image<char> lap(width, height);
    Image<char> lap_out(width, height);
    Image<char> restored(width, height);
    Mask<float> mask(coef);

    Pyramid<char> pg(gauss, depth);
    Pyramid<char> pb(blurred, depth);
    Pyramid<char> pl(lap, depth);

    traverse(pg, pb, pl, [&] () {
        if (!pg.is_top_level()) {
            BoundaryCondition<char> bc(pg(-1), mask, Boundary::MIRROR);
            Accessor<char> a_fine(bc);
            IterationSpace<char> i_blur(pb(-1));
            Blur blur(i_blur, a_fine, mask);
            blur.execute();

            Accessor<char> a_blur(pb(-1), Interpolate::NN);
            IterationSpace<char> i_down(pg(0));
            Take down(i_down, a_blur);
            down.execute();

            Accessor<char> a_prev(pg(-1));
            Accessor<char> a_up(pg(0), Interpolate::LF);
            IterationSpace<char> i_lap(pl(-1));
            Detail detail(i_lap, a_prev, a_up);
            detail.execute();
        }

        traverse();

        if (!pg.is_bottom_level()) {
            Accessor<char> a_coarse(pg(1), Interpolate::LF);
            Accessor<char> a_detail(pl(0));
            IterationSpace<char> i_back(pg(0));
            Rebuild rebuild(i_back, a_coarse, a_detail);
            rebuild.execute();
        }

        if (pg.is_top_level()) {
            Accessor<char> a_g0(pg(0));
            IterationSpace<char> i_restored(restored);
            Take keep_g(i_restored, a_g0);
            keep_g.execute();

            Accessor<char> a_l0(pl(0));
            IterationSpace<char> i_lap_out(lap_out);
            Take keep_l(i_lap_out, a_l0);
            keep_l.execute();
        }
    });

    Image<short> work(width, height, start);
    Image<short> next(width, height);
    Accessor<short> a_work(work);
    IterationSpace<short> i_next(next);
    for (int i = 0; i < 3; ++i) {
        // a new instance per run: the DSL host emulation executes an
        // instance once (dsl/kernel.hpp:94), compiled code at every execute()
        Step step(i_next, a_work);
        step.execute();
        work = next;
    }

    Domain dom(3, 3);
    dom(0, 0) = 0;
    BoundaryCondition<char> bc_ring(restored, dom, Boundary::REPEAT);
    Accessor<char> a_ring(bc_ring);
    Image<char> ringmax(width, height);
    IterationSpace<char> i_ring(ringmax);
    RingMax ring(i_ring, a_ring, dom);
    ring.execute();

    FILE *f = std::fopen(argv[1], "wb");
    std::fwrite(lap_out.data(), 1, width * height, f);
    std::fwrite(restored.data(), 1, width * height, f);
    std::fwrite(work.data(), sizeof(short), width * height, f);
    std::fwrite(ringmax.data(), 1, width * height, f);
    std::fclose(f);
    delete[] pixels;
    delete[] start;
    return 0;
}
