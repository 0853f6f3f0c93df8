"""The oracle's pyramid (tests/oracle_binding.pyramid, the pyramid sample's
per-level kernel sequence over the C oracle) against what fixes it:
level sizes of dsl/pyramid.hpp:103-114, the sample's own round-trip check
(samples-public/5_Other/Gaussian_Laplacian_Pyramid/src/main.cpp:228-260:
restoring G[l] = LF(G[l+1]) + L[l] from the coarsest level returns the
input), the resampling of a constant image, and the NN map for even sizes."""
import numpy as np
import pytest

import oracle_binding as ob
from hipacc_amd import capi


@pytest.mark.parametrize("shape,levels", [((29, 37), 4), ((64, 48), 5), ((17, 130), 4)])
def test_pyramid_levels_and_round_trip(shape, levels):
    g0 = ob.random_image(shape, np.float32, 77)
    G, L = ob.pyramid(g0, levels)
    for l, g in enumerate(G):
        assert g.shape == (shape[0] >> l, shape[1] >> l)
    # collapse: G[l] = LF(G[l+1]) + L[l], from the coarsest level up
    r = G[-1]
    for l in range(levels - 2, -1, -1):
        out = np.zeros_like(G[l])
        ob.resample(r, out, capi.INTERP_LF, op=capi.RESAMPLE_ADD, src2=L[l])
        # (g - v) + v returns g up to one rounding of the larger operand
        assert np.max(np.abs(out - G[l])) <= 2.0 ** -22
        r = G[l]
    # full collapse from the coarsest level alone plus the Laplacians
    r = G[-1]
    for l in range(levels - 2, -1, -1):
        out = np.zeros_like(G[l])
        ob.resample(r, out, capi.INTERP_LF, op=capi.RESAMPLE_ADD, src2=L[l])
        r = out
    assert np.max(np.abs(r - g0)) <= 2.0 ** -20


def test_lf_and_nn_of_constant_image():
    src = np.full((13, 22), 0.375, np.float32)
    for dst_shape in ((6, 11), (27, 45)):
        for interp in (capi.INTERP_NN, capi.INTERP_LF):
            dst = np.zeros(dst_shape, np.float32)
            ob.resample(src, dst, interp)
            if interp == capi.INTERP_NN:
                assert np.all(dst == np.float32(0.375))
            else:  # the four float weights sum to 1 up to rounding
                assert np.max(np.abs(dst - np.float32(0.375))) <= 4 * 2.0 ** -24


def test_nn_downsample_of_even_size_picks_odd_pixels():
    """For even sizes the NN map is x -> 2x+1 (s = 2: s/2 + 2x)."""
    src = np.arange(16 * 12, dtype=np.float32).reshape(12, 16)
    dst = np.zeros((6, 8), np.float32)
    ob.resample(src, dst, capi.INTERP_NN)
    np.testing.assert_array_equal(dst, src[1::2, 1::2])


# ---- the plan's work lists (host only) -------------------------------------
def _bench_shapes():
    import bench
    return bench.pyramid_workload_shapes()


@pytest.mark.parametrize("tile_rows", [8, 16, 32])
def test_plan_tiles_fit_on_benchmark_geometry(tile_rows):
    """Every tile of the 64-image, 8-level benchmark pyramid fits the LDS
    capacity of the kernel that runs it (the library proves it on the host
    and fails the plan otherwise), and levels 0-3 run on wide tiles."""
    from hipacc_amd import operators as op
    narrow, wide = op.pyramid_plan_tiles(_bench_shapes(), 8, tile_rows)
    assert all(w >= 512 for w in wide[:4]), wide
    assert all(w == 0 for w in wide[4:]), wide
    shapes = _bench_shapes()
    for l in range(7):
        # every pixel of level l is covered exactly once by the two lists
        if wide[l]:
            assert wide[l] == sum(-(-(h >> l) // 32) * -(-(w >> l) // 128) for h, w in shapes
                                  if (h >> l) >= 9 and (w >> l) >= 9)
        else:
            assert narrow[l] == sum(-(-(h >> l) // tile_rows) * -(-(w >> l) // 64)
                                    for h, w in shapes)


def test_plan_tiles_mixed_threshold_geometry():
    """The GPU test geometry: two 2048^2 images fill level 1 with 512 wide
    tiles, next to images of 8 and 9 rows at that level."""
    from hipacc_amd import operators as op
    import test_gpu_pyramid_wide as tw
    narrow, wide = op.pyramid_plan_tiles(tw.MIXED, tw.MIXED_LEVELS)
    assert wide[0] >= 512 and wide[1] >= 512, wide
    assert narrow[1] > 0, narrow   # the 8-row image of level 1 takes narrow tiles


def test_plan_tiles_rejects_bad_sizes():
    from hipacc_amd import operators as op
    with pytest.raises(capi.HipaccError):
        op.pyramid_plan_tiles([(3, 300)], 4)   # no pixel left at level 2


# ---- the interpolation set (NN / LF / B5 / CF / L3) ---------------------------
def test_interpolation_closed_forms():
    """B5 of a constant image is c * Sx * Sy, S = 11/8 (1 where frac is 0
    or 0.5); CF with the reference's tap weights scales a constant by the
    sums of its x and y weights; NN of an up-sampled image repeats source
    pixels."""
    c = np.full((20, 24), 0.75, np.float32)
    out = np.zeros((33, 41), np.float32)
    ob.resample(c, out, capi.INTERP_B5, boundary=capi.CLAMP)

    def tap_sum(n_src, n_dst):
        # the binomial weights at distances frac + 0.5 - k sum to 11/8, or to
        # 1 where a distance hits a step of the weight (frac 0 or 0.5)
        s = np.float32(n_src) / np.float32(n_dst)
        xm = (s / np.float32(2)) + s * np.arange(n_dst, dtype=np.float32)
        xb = np.maximum(xm - np.float32(0.5), np.float32(0))
        f = xb - np.floor(xb)
        return np.where((f == 0) | (f == 0.5), 1.0, 11 / 8)
    want = 0.75 * np.outer(tap_sum(20, 33), tap_sum(24, 41))
    np.testing.assert_allclose(out, want, rtol=1e-6)
    # CF: pixel x_int - 1 + k takes Keys' weight of (frac - 1) + k, as the
    # reference does (dsl/image.hpp:455-478 and the generated CUDA code):
    # a constant image gives c * Sx * Sy with S = sum_k w(frac - 1 + k),
    # which is not 1 (the weights are mirrored about the pixel)
    ob.resample(c, out, capi.INTERP_CF, boundary=capi.CLAMP)

    def keys(d):
        d = np.abs(d)
        return np.where(d < 1, 1.5 * d ** 3 - 2.5 * d ** 2 + 1,
                        np.where(d < 2, -0.5 * d ** 3 + 2.5 * d ** 2 - 4 * d + 2, 0.0))

    def cf_sum(n_src, n_dst):
        s = np.float32(n_src) / np.float32(n_dst)
        xm = (s / np.float32(2)) + s * np.arange(n_dst, dtype=np.float32)
        xb = np.maximum(xm - np.float32(0.5), np.float32(0)).astype(np.float64)
        f = xb - np.floor(xb)
        return sum(keys(f - 1 + k) for k in range(4))
    want = 0.75 * np.outer(cf_sum(20, 33), cf_sum(24, 41))
    np.testing.assert_allclose(out, want, rtol=1e-5)
    assert np.abs(want - 0.75).max() > 0.05   # the mirrored weights do not sum to 1
    y, x = np.mgrid[0:20, 0:24].astype(np.float32)
    ramp = (0.25 * x + 0.5 * y).astype(np.float32)
    nn = np.zeros((40, 48), np.float32)
    ob.resample(ramp, nn, capi.INTERP_NN, boundary=capi.CLAMP)
    np.testing.assert_array_equal(nn, np.repeat(np.repeat(ramp, 2, 0), 2, 1))


def test_interpolation_l3_reference_rows():
    """L3 reads rows y_int - 1 .. y_int + 4 weighted lanczos(y_frac - 2 + k),
    and the second row's fifth tap takes the sixth tap's weight, as the
    reference (dsl/image.hpp:455-532): an image that is zero except one row
    shows which source rows feed an output row."""
    src = np.zeros((16, 16), np.float32)
    src[8, :] = 1.0
    out = np.zeros((16, 16), np.float32)
    ob.resample(src, out, capi.INTERP_L3, boundary=capi.CLAMP)
    # same size: frac = 0; row y reads rows y - 1 .. y + 4 with weights
    # l3(-2), l3(-1), l3(0) = 1, l3(1), l3(2), l3(3) = 0 -> only row y+1
    # contributes (weight 1) among the rows with nonzero weight at integer
    # distances, so output row 7 holds the source row 8
    np.testing.assert_allclose(out[7], 1.0, atol=1e-6)
    assert np.all(np.abs(np.delete(out, 7, axis=0)) < 1e-6)
