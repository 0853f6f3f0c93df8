// Emitted by hipacc_amd.emit from pyramid_i8.cpp for the MI355X runtime (hipacc_amd/runtime/hipacc_hip.hpp).
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
#include "hipacc_hip_device.hpp"
// DSL names kept by the emitted host code
typedef unsigned char uchar;
typedef unsigned short ushort;
typedef unsigned int uint;
#ifndef HIPACC_CODEGEN
#define HIPACC_CODEGEN
#endif
namespace hipacc { namespace math { using std::min; using std::max; } }
namespace hipacc { using Reduce = hipacc_dev::Reduce; }

#include <cstdint>
#include <cstdio>

using namespace hipacc;
using namespace hipacc::math;

// Kernel class Blur: local operator SUM over mask mask (hipaccLaunchLocal)

// Kernel class Take: device code, the kernel() body compiled as written (hipaccLaunchUser)
struct Take : hipacc_dev::Kernel<char> {
    hipacc_dev::Accessor<char, true> in;
    __device__ void set_pos(int x_, int y_) { it_x = x_; it_y = y_; in.px = x_; in.py = y_; }
    __device__ __forceinline__ void kernel() { output() = in(); }
};

// Kernel class Detail: device code, the kernel() body compiled as written (hipaccLaunchUser)
struct Detail : hipacc_dev::Kernel<char> {
    hipacc_dev::Accessor<char> fine;
    hipacc_dev::Accessor<char, true> coarse;
    __device__ void set_pos(int x_, int y_) { it_x = x_; it_y = y_; fine.px = x_; fine.py = y_; coarse.px = x_; coarse.py = y_; }
    __device__ __forceinline__ void kernel() { output() = fine() - coarse(); }
};

// Kernel class Rebuild: device code, the kernel() body compiled as written (hipaccLaunchUser)
struct Rebuild : hipacc_dev::Kernel<char> {
    hipacc_dev::Accessor<char, true> coarse;
    hipacc_dev::Accessor<char> detail;
    __device__ void set_pos(int x_, int y_) { it_x = x_; it_y = y_; coarse.px = x_; coarse.py = y_; detail.px = x_; detail.py = y_; }
    __device__ __forceinline__ void kernel() { output() = coarse() + detail(); }
};

// Kernel class Step: device code, the kernel() body compiled as written (hipaccLaunchUser)
struct Step : hipacc_dev::Kernel<short> {
    hipacc_dev::Accessor<short> in;
    __device__ void set_pos(int x_, int y_) { it_x = x_; it_y = y_; in.px = x_; in.py = y_; }
    __device__ __forceinline__ void kernel() {
        short v = in();
        output() = v > 0 ? v / 2 : v + 11;
    }
};

// Kernel class RingMax: local operator MAX over domain dom (hipaccLaunchLocal)

static uint32_t mix(int x, int y, uint32_t salt) {
    uint32_t h = (uint32_t)x * 2654435761u ^ (uint32_t)y * 2246822519u ^ salt;
    h = (h ^ (h >> 16)) * 2146121005u;
    return h ^ (h >> 15);
}

HIPACC_CODEGEN int main(int argc, const char **argv) {
    hipaccInitHIP(0);
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

    HipaccImageHip<char> gauss = hipaccCreateMemory<char>(pixels, width, height);
    HipaccImageHip<char> blurred = hipaccCreateMemory<char>(nullptr, width, height);
    HipaccImageHip<char> lap = hipaccCreateMemory<char>(nullptr, width, height);
    HipaccImageHip<char> lap_out = hipaccCreateMemory<char>(nullptr, width, height);
    HipaccImageHip<char> restored = hipaccCreateMemory<char>(nullptr, width, height);
    HipaccMask<float> mask(5, 5, &coef[0][0]);

    HipaccPyramidHip<char> pg = hipaccCreatePyramid<char>(gauss, depth);
    HipaccPyramidHip<char> pb = hipaccCreatePyramid<char>(blurred, depth);
    HipaccPyramidHip<char> pl = hipaccCreatePyramid<char>(lap, depth);

    HipaccPyramidTraversor hipaccTraversor;
    hipaccTraversor.hipaccTraverse(pg, pb, pl, [&] () {
        if (!pg.is_top_level()) {
            // BoundaryCondition bc: MIRROR of pg(-1)
            HipaccAccessor<char> a_fine = hipaccMakeAccessor(pg(-1));
            HipaccAccessor<char> i_blur = hipaccMakeAccessor(pb(-1));
            // Blur blur(i_blur, a_fine, mask): launched at blur.execute()
            {
                HipaccLocalOptions blur_opt;
                blur_opt.reduce = HipaccReduce::SUM;
                blur_opt.boundary = HipaccBoundary::MIRROR;
                blur_opt.use_weights = true;
                blur_opt.fp_contract = false;
                auto blur_mask = hipaccMaskCast<float>(mask);
                hipaccLaunchLocal(a_fine, i_blur, blur_mask, blur_opt);
            }

            HipaccAccessor<char> a_blur = hipaccMakeAccessor(pb(-1));
            HipaccAccessor<char> i_down = hipaccMakeAccessor(pg(0));
            Take down;
            down.in = hipaccDevAccessor<true>(a_blur, HipaccBoundary::UNDEFINED, 0, HIPACC_INTERP_NN, &i_down);
            hipaccLaunchUser(down, i_down);

            HipaccAccessor<char> a_prev = hipaccMakeAccessor(pg(-1));
            HipaccAccessor<char> a_up = hipaccMakeAccessor(pg(0));
            HipaccAccessor<char> i_lap = hipaccMakeAccessor(pl(-1));
            Detail detail;
            detail.fine = hipaccDevAccessor(a_prev, HipaccBoundary::UNDEFINED, 0);
            detail.coarse = hipaccDevAccessor<true>(a_up, HipaccBoundary::UNDEFINED, 0, HIPACC_INTERP_LF, &i_lap);
            hipaccLaunchUser(detail, i_lap);
        }

        hipaccTraversor.hipaccTraverse();

        if (!pg.is_bottom_level()) {
            HipaccAccessor<char> a_coarse = hipaccMakeAccessor(pg(1));
            HipaccAccessor<char> a_detail = hipaccMakeAccessor(pl(0));
            HipaccAccessor<char> i_back = hipaccMakeAccessor(pg(0));
            Rebuild rebuild;
            rebuild.coarse = hipaccDevAccessor<true>(a_coarse, HipaccBoundary::UNDEFINED, 0, HIPACC_INTERP_LF, &i_back);
            rebuild.detail = hipaccDevAccessor(a_detail, HipaccBoundary::UNDEFINED, 0);
            hipaccLaunchUser(rebuild, i_back);
        }

        if (pg.is_top_level()) {
            HipaccAccessor<char> a_g0 = hipaccMakeAccessor(pg(0));
            HipaccAccessor<char> i_restored = hipaccMakeAccessor(restored);
            Take keep_g;
            keep_g.in = hipaccDevAccessor<true>(a_g0, HipaccBoundary::UNDEFINED, 0);
            hipaccLaunchUser(keep_g, i_restored);

            HipaccAccessor<char> a_l0 = hipaccMakeAccessor(pl(0));
            HipaccAccessor<char> i_lap_out = hipaccMakeAccessor(lap_out);
            Take keep_l;
            keep_l.in = hipaccDevAccessor<true>(a_l0, HipaccBoundary::UNDEFINED, 0);
            hipaccLaunchUser(keep_l, i_lap_out);
        }
    });

    HipaccImageHip<short> work = hipaccCreateMemory<short>(start, width, height);
    HipaccImageHip<short> next = hipaccCreateMemory<short>(nullptr, width, height);
    HipaccAccessor<short> a_work = hipaccMakeAccessor(work);
    HipaccAccessor<short> i_next = hipaccMakeAccessor(next);
    for (int i = 0; i < 3; ++i) {
        // a new instance per run: the DSL host emulation executes an
        // instance once (dsl/kernel.hpp:94), compiled code at every execute()
        Step step;
        step.in = hipaccDevAccessor(a_work, HipaccBoundary::UNDEFINED, 0);
        hipaccLaunchUser(step, i_next);
        hipaccCopyMemory(next, work);
    }

    HipaccMask<float> dom(3, 3);
    dom.domain_at(0,0) = 0;
    // BoundaryCondition bc_ring: REPEAT of restored
    HipaccAccessor<char> a_ring = hipaccMakeAccessor(restored);
    HipaccImageHip<char> ringmax = hipaccCreateMemory<char>(nullptr, width, height);
    HipaccAccessor<char> i_ring = hipaccMakeAccessor(ringmax);
    // RingMax ring(i_ring, a_ring, dom): launched at ring.execute()
    {
        HipaccLocalOptions ring_opt;
        ring_opt.reduce = HipaccReduce::MAX;
        ring_opt.boundary = HipaccBoundary::REPEAT;
        ring_opt.use_weights = false;
        ring_opt.fp_contract = false;
        auto ring_mask = hipaccMaskCast<int>(dom);
        hipaccLaunchLocal(a_ring, i_ring, ring_mask, ring_opt);
    }

    FILE *f = std::fopen(argv[1], "wb");
    std::fwrite(hipaccReadMemory<char>(lap_out), 1, width * height, f);
    std::fwrite(hipaccReadMemory<char>(restored), 1, width * height, f);
    std::fwrite(hipaccReadMemory<short>(work), sizeof(short), width * height, f);
    std::fwrite(hipaccReadMemory<char>(ringmax), 1, width * height, f);
    std::fclose(f);
    delete[] pixels;
    delete[] start;
    return 0;
}
