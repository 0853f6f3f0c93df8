"""The 8-bit block row ring (16 pixels per lane, balanced strips) and the
separable ring with integer sums of 8-bit pixels: interior
strips staged by direct loads, edge strips with per-pixel loads, every
boundary mode, windows with offsets, pitched and unaligned rows, batches,
segment heights, float sums and exact integer sums -- bit for bit against
the oracle, with the ring variant forced (an unsupported shape fails the
test instead of falling back)."""
import numpy as np
import pytest

from hipacc_amd import capi, masks, operators as op

import oracle_binding as ob
from gpu_util import empty_gpu, to_gpu, to_host

pytestmark = pytest.mark.gpu

MODES = [capi.UNDEFINED, capi.CLAMP, capi.REPEAT, capi.MIRROR, capi.CONSTANT]
# (shape, pitch): 4 balanced strips with 2 interior; a full-width 4096 row
# with padding; a ragged right edge; rows not 16-byte aligned (edge path
# only); a batch of planes
SHAPES = [((37, 3000), 0), ((23, 4096), 4112), ((19, 2085), 2096), ((21, 1501), 1501),
          ((2, 17, 2500), 0)]


def _ring(src, out_dt, *, pitch=0, seg=0, src_window=None, dst_window=None, dst_shape=None,
          **kw):
    shape = dst_shape if dst_shape is not None else src.shape
    ref = np.zeros(shape, out_dt)
    ob.local(src, ref, src_window=src_window, dst_window=dst_window, **kw)
    dst = empty_gpu(shape, out_dt, pitch=pitch)
    call = op.prepare_local(to_gpu(src, pitch=pitch), dst, src_window=src_window,
                            dst_window=dst_window, dst_dtype=ob.CODES[np.dtype(out_dt)], **kw)
    call.with_variant(capi.VAR_RING, seg).launch()
    return to_host(dst, out_dt), ref


def _cmp(got, ref, mode, r, msg):
    if mode == capi.UNDEFINED:
        got, ref = got[..., r:-r, r:-r], ref[..., r:-r, r:-r]
    np.testing.assert_array_equal(got, ref, err_msg=msg)


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("k", [3, 5])
@pytest.mark.parametrize("shape,pitch", SHAPES)
def test_u8_gauss_ring(gpu, mode, k, shape, pitch):
    print(f"ring8 gauss k={k} mode={mode} shape={shape} pitch={pitch}", flush=True)
    src = ob.random_image(shape, np.uint8, 7 * k + shape[-1])
    w = masks.gaussian_2d(k, k / 4.0)
    for contract, cvt in ((True, capi.CVT_ROUND), (False, capi.CVT_TRUNC)):
        got, ref = _ring(src, np.uint8, pitch=pitch, mask_w=k, mask_h=k, weights=w,
                         boundary=mode, const_val=201, cvt=cvt, fp_contract=contract)
        _cmp(got, ref, mode, k // 2, f"contract={contract} cvt={cvt}")


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("out", [np.int16, np.uint32, np.int32])
@pytest.mark.parametrize("shape,pitch", SHAPES[:4])
def test_u8_sobel_ring(gpu, mode, out, shape, pitch):
    print(f"ring8 sobel out={np.dtype(out)} mode={mode} shape={shape}", flush=True)
    src = ob.random_image(shape, np.uint8, 3 + shape[-1])
    for m in (masks.SOBEL_X, masks.SOBEL_Y):
        dom = (m != 0).astype(np.uint8)
        got, ref = _ring(src, out, pitch=pitch, mask_w=3, mask_h=3, weights=m, domain=dom,
                         acc_float=False, boundary=mode, const_val=17)
        _cmp(got, ref, mode, 1, "sobel")


@pytest.mark.parametrize("mode", MODES)
def test_u8_int5x5_ring(gpu, mode):
    """A 5x5 integer mask with mixed signs whose sums reach about 2^23 (still
    exact in float): wrap-free int arithmetic, i32 output."""
    src = ob.random_image((29, 3100), np.uint8, 41)
    w = np.random.default_rng(5).integers(-2600, 2600, 25).astype(np.int32)
    got, ref = _ring(src, np.int32, mask_w=5, mask_h=5, weights=w, acc_float=False,
                     boundary=mode, const_val=250)
    _cmp(got, ref, mode, 2, "int 5x5")


@pytest.mark.parametrize("seg", [1, 5, 8, 16, 37])
def test_u8_ring_segments(gpu, seg):
    """Forced segment heights: partial chunks, segments shorter than the mask."""
    src = ob.random_image((83, 3000), np.uint8, seg)
    w = masks.gaussian_2d(5, 1.0)
    for mode in (capi.CLAMP, capi.CONSTANT):
        got, ref = _ring(src, np.uint8, seg=seg, mask_w=5, mask_h=5, weights=w, boundary=mode,
                         const_val=9, cvt=capi.CVT_ROUND)
        _cmp(got, ref, mode, 2, f"seg={seg} mode={mode}")


@pytest.mark.parametrize("mode", [capi.CLAMP, capi.MIRROR, capi.REPEAT, capi.CONSTANT])
def test_u8_ring_windows(gpu, mode):
    """Accessor windows and iteration spaces with offsets: 16-aligned offsets
    keep the direct-load path, others take the edge path."""
    src = ob.random_image((61, 3333), np.uint8, 99)
    w = masks.gaussian_2d(5, 1.0)
    for sw, dw, pitch in (((32, 2, 3200, 55), (16, 4, 3200, 55), 3344),
                          ((8, 0, 3000, 61), (0, 0, 3000, 61), 3344),
                          ((0, 5, 3333, 50), (0, 0, 3333, 50), 3340)):
        got, ref = _ring(src, np.uint8, pitch=pitch, mask_w=5, mask_h=5, weights=w, boundary=mode,
                         const_val=3, cvt=capi.CVT_ROUND, src_window=sw, dst_window=dw)
        np.testing.assert_array_equal(got, ref, err_msg=f"window {sw} -> {dw}")


def test_u8_ring_refuses_inexact_int_masks(gpu):
    """Integer masks whose sums could leave float's exact range are not taken
    by the ring; the automatic choice still gives the oracle's result."""
    src = ob.random_image((20, 3000), np.uint8, 1)
    w = np.full(9, 1 << 16, np.int32)
    ref = np.zeros(src.shape, np.int32)
    ob.local(src, ref, 3, 3, weights=w, acc_float=False, boundary=capi.CLAMP)
    dst = empty_gpu(src.shape, np.int32)
    call = op.prepare_local(to_gpu(src), dst, 3, 3, weights=w, acc_float=False,
                            boundary=capi.CLAMP, dst_dtype=capi.I32)
    with pytest.raises(capi.HipaccError) as e:
        call.with_variant(capi.VAR_RING, 0).launch()
    assert e.value.status == 2
    call.launch()
    np.testing.assert_array_equal(to_host(dst, np.int32), ref)


# ---- separable ring with integer sums ------------------------------------------
def _ring_sep(src, out_dt, row, col, *, pitch=0, seg=0, **kw):
    ref = np.zeros(src.shape, out_dt)
    ob.separable(src, ref, row, col, acc_float=False, dst_dtype=ob.CODES[np.dtype(out_dt)],
                 **kw)
    dst = empty_gpu(src.shape, out_dt, pitch=pitch)
    call = op.prepare_separable(to_gpu(src, pitch=pitch), dst, row, col, acc_float=False,
                                dst_dtype=ob.CODES[np.dtype(out_dt)], **kw)
    call.with_variant(capi.VAR_RING, seg).launch()
    return to_host(dst, out_dt), ref


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("k", [3, 5, 7])
@pytest.mark.parametrize("out", [np.uint32, np.int32, np.int16])
def test_u8_box_sep_ring(gpu, mode, k, out):
    """Box sums (all-ones rows and columns) of u8 pixels, exact in float."""
    print(f"ring sep box k={k} out={np.dtype(out)} mode={mode}", flush=True)
    src = ob.random_image((45, 2100), np.uint8, 11 * k)
    ones = np.ones(k, np.int32)
    for pitch in (0, 2112):
        got, ref = _ring_sep(src, out, ones, ones, pitch=pitch, boundary=mode, const_val=77)
        _cmp(got, ref, mode, k // 2, f"pitch={pitch}")


@pytest.mark.parametrize("mode", MODES)
def test_u8_int_sep_ring(gpu, mode):
    """Mixed-sign integer rows and columns (a separable Sobel and a wider
    derivative), i16 and i32 outputs; a negative CONSTANT value, which the
    pixel type wraps."""
    src = ob.random_image((39, 1700), np.uint8, 23)
    for row, col, out in (([1, 2, 1], [-1, 0, 1], np.int16),
                          ([1, 4, 6, 4, 1], [-2, -1, 0, 1, 2], np.int32),
                          ([3, -7, 11, 5, -2, 9, 1], [1, 1, 2, 3, 2, 1, 1], np.int32)):
        got, ref = _ring_sep(src, out, np.array(row, np.int32), np.array(col, np.int32),
                             boundary=mode, const_val=-900)
        _cmp(got, ref, mode, len(row) // 2, f"row={row}")


def test_u8_sep_ring_refuses_inexact_masks(gpu):
    src = ob.random_image((20, 600), np.uint8, 2)
    row = np.full(7, 300, np.int32)
    call = op.prepare_separable(to_gpu(src), empty_gpu(src.shape, np.int32), row, row,
                                acc_float=False, dst_dtype=capi.I32)
    with pytest.raises(capi.HipaccError) as e:
        call.with_variant(capi.VAR_RING, 0).launch()
    assert e.value.status == 2
