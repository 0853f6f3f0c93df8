"""The image-processing operators the benchmark times, at their full 4096x4096
size and through the automatic kernel choice (the geometry the benchmark
runs: balanced u8 strips, segment heights sized to one round of resident
blocks), bit for bit against the oracle."""
import numpy as np
import pytest

from hipacc_amd import capi, masks, operators as op

import oracle_binding as ob
from gpu_util import empty_gpu, to_gpu, to_host
from weights import skewed_1d

pytestmark = pytest.mark.gpu

N = 4096


@pytest.fixture(scope="module")
def u8_image():
    return ob.random_image((N, N), np.uint8, 4096)


@pytest.fixture(scope="module")
def f32_image():
    return ob.random_image((N, N), np.float32, 4097)


@pytest.mark.parametrize("mode", [capi.CLAMP, capi.MIRROR])
def test_sobel_u8_i16_full(gpu, u8_image, mode):
    print(f"full-size Sobel u8 -> i16 mode={mode}", flush=True)
    sob = masks.SOBEL_X
    dom = (sob != 0).astype(np.uint8)
    ref = np.zeros((N, N), np.int16)
    ob.local(u8_image, ref, 3, 3, weights=sob, domain=dom, acc_float=False, boundary=mode)
    dst = empty_gpu((N, N), np.int16)
    op.local(to_gpu(u8_image), dst, 3, 3, weights=sob, domain=dom, acc_float=False,
             boundary=mode, dst_dtype=capi.I16)
    np.testing.assert_array_equal(to_host(dst, np.int16), ref)


def test_sobel_f32_full(gpu, f32_image):
    print("full-size Sobel f32", flush=True)
    sob = masks.SOBEL_X
    ref = np.zeros((N, N), np.float32)
    ob.local(f32_image, ref, 3, 3, weights=sob, boundary=capi.CLAMP)
    dst = empty_gpu((N, N), np.float32)
    op.local(to_gpu(f32_image), dst, 3, 3, weights=sob, boundary=capi.CLAMP)
    np.testing.assert_array_equal(to_host(dst, np.float32), ref)


@pytest.mark.parametrize("dt", [np.float32, np.uint8])
@pytest.mark.parametrize("skew", [False, True])
def test_gauss13_separable_full(gpu, u8_image, f32_image, dt, skew):
    """The benchmarked 13-tap Gaussian, and (skew) the same row with a
    different column, which pins the orientation of the two passes."""
    print(f"full-size 13-tap separable {np.dtype(dt)} skew={skew}", flush=True)
    src = f32_image if dt == np.float32 else u8_image
    g13 = masks.gaussian_1d(13, 3.0)
    c13 = skewed_1d(13, 1313) if skew else g13
    cvt = capi.CVT_ROUND if dt == np.uint8 else capi.CVT_TRUNC
    ref = np.zeros((N, N), dt)
    ob.separable(src, ref, g13, c13, boundary=capi.CLAMP, cvt=cvt)
    dst = empty_gpu((N, N), dt)
    op.prepare_separable(to_gpu(src), dst, g13, c13, boundary=capi.CLAMP, cvt=cvt).launch()
    np.testing.assert_array_equal(to_host(dst, dt), ref)


def test_box7_u8_u32_full(gpu, u8_image):
    print("full-size 7x7 box u8 -> u32", flush=True)
    ones = np.ones(7, np.int32)
    ref = np.zeros((N, N), np.uint32)
    ob.separable(u8_image, ref, ones, ones, acc_float=False, boundary=capi.CLAMP,
                 dst_dtype=capi.U32)
    dst = empty_gpu((N, N), np.uint32)
    op.prepare_separable(to_gpu(u8_image), dst, ones, ones, acc_float=False,
                         boundary=capi.CLAMP, dst_dtype=capi.U32).launch()
    np.testing.assert_array_equal(to_host(dst, np.uint32), ref)


def test_gauss5_u8_full(gpu, u8_image):
    print("full-size 5x5 u8 rounded", flush=True)
    w = masks.gaussian_2d(5, 1.0)
    ref = np.zeros((N, N), np.uint8)
    ob.local(u8_image, ref, 5, 5, weights=w, boundary=capi.CLAMP, cvt=capi.CVT_ROUND)
    dst = empty_gpu((N, N), np.uint8)
    op.local(to_gpu(u8_image), dst, 5, 5, weights=w, boundary=capi.CLAMP, cvt=capi.CVT_ROUND)
    np.testing.assert_array_equal(to_host(dst, np.uint8), ref)
