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
