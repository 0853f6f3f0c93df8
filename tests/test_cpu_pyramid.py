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
