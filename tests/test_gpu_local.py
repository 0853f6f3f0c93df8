"""GPU parity of the local operators (libhipacc_gfx950.so through the C ABI)
against the reference DSL golden vectors and the C oracle.

Bar: bit-exact for every integer kernel and for float kernels evaluated in the
oracle's rounding order (fp_contract=0: mul then add, as the DSL host
emulation; fp_contract=1: fmaf, as the reference CUDA backend with nvcc's
default contraction).  UNDEFINED is compared on the interior only (the
reference reads outside the image there).
"""
import numpy as np
import pytest
import torch

from hipacc_amd import capi, masks, operators as op

import golden_cases as gc
import oracle_binding as ob
from gpu_util import empty_gpu, to_gpu, to_host

pytestmark = pytest.mark.gpu

MODES = [capi.UNDEFINED, capi.CLAMP, capi.REPEAT, capi.MIRROR, capi.CONSTANT]
GOLDEN = [(c, e) for c, e in gc.load() if gc.local_spec(c) is not None]


@pytest.mark.parametrize("case,expected", GOLDEN, ids=[gc.case_id(c) for c, _ in GOLDEN])
def test_local_matches_reference_dsl(gpu, case, expected):
    print(f"golden {gc.case_id(case)}", flush=True)
    spec = gc.local_spec(case)
    out_dtype = spec.pop("out_dtype")
    src = gc.input_for(case)
    dst = empty_gpu(expected.shape, out_dtype)
    spec.setdefault("dst_dtype", ob.CODES[np.dtype(out_dtype)])
    op.local(to_gpu(src), dst, **spec)
    np.testing.assert_array_equal(to_host(dst, out_dtype), expected)


def _run_both(src, out_dtype, *, pitch=0, src_window=None, dst_window=None, dst_shape=None, **kw):
    shape = dst_shape if dst_shape is not None else src.shape
    ref = np.zeros(shape, dtype=out_dtype)
    ob.local(src, ref, src_window=src_window, dst_window=dst_window, **kw)
    dst = empty_gpu(shape, out_dtype, pitch=pitch)
    op.local(to_gpu(src, pitch=pitch), dst, src_window=src_window, dst_window=dst_window,
             dst_dtype=ob.CODES[np.dtype(out_dtype)], **kw)
    return to_host(dst, out_dtype), ref


def _interior(a, r):
    return a[..., r:a.shape[-2] - r, r:a.shape[-1] - r]


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("k", [3, 5, 7, 9, 13])
@pytest.mark.parametrize("contract", [False, True])
def test_f32_gaussian_vs_oracle(gpu, mode, k, contract):
    print(f"f32 gaussian k={k} mode={mode} contract={contract}", flush=True)
    src = ob.random_image((70, 300), np.float32, 11 + k)
    got, ref = _run_both(src, np.float32, mask_w=k, mask_h=k, weights=masks.gaussian_2d(k, k / 4),
                         boundary=mode, const_val=0.25, fp_contract=contract)
    if mode == capi.UNDEFINED:
        got, ref = _interior(got, k // 2), _interior(ref, k // 2)
    np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("k", [3, 5, 13])
def test_u8_gaussian_vs_oracle(gpu, mode, k):
    print(f"u8 gaussian k={k} mode={mode}", flush=True)
    src = ob.random_image((45, 1100), np.uint8, 7 + k)
    for contract in (False, True):
        got, ref = _run_both(src, np.uint8, mask_w=k, mask_h=k, weights=masks.gaussian_2d(k, k / 4),
                             boundary=mode, const_val=200, cvt=capi.CVT_ROUND, fp_contract=contract)
        if mode == capi.UNDEFINED:
            got, ref = _interior(got, k // 2), _interior(ref, k // 2)
        np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("out", [np.int16, np.uint32, np.int32])
@pytest.mark.parametrize("k", [3, 5, 7])
def test_int_masks_vs_oracle(gpu, mode, out, k):
    print(f"int mask k={k} out={out} mode={mode}", flush=True)
    src = ob.random_image((33, 1500), np.uint8, 3 + k)
    rng = np.random.default_rng(k)
    w = rng.integers(-9, 10, size=(k, k)).astype(np.int32)
    dom = (rng.random((k, k)) > 0.3).astype(np.uint8)
    dom[k // 2, k // 2] = 1
    for weights, domain in ((w, None), (w, dom), (None, None), (None, dom)):
        got, ref = _run_both(src, out, mask_w=k, mask_h=k, weights=weights, domain=domain,
                             acc_float=False, boundary=mode, const_val=17)
        if mode == capi.UNDEFINED:
            got, ref = _interior(got, k // 2), _interior(ref, k // 2)
        np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("red", [capi.MIN, capi.MAX, capi.PROD, capi.SUM])
def test_reduce_modes_generic_vs_oracle(gpu, mode, red):
    print(f"reduce mode={red} boundary={mode}", flush=True)
    k = 5
    src = ob.random_image((40, 90), np.float32, 5)
    dom = np.ones((k, k), np.uint8)
    dom[0, 0] = dom[4, 4] = dom[2, 3] = 0
    for weights in (None, masks.gaussian_2d(k, 1.0) * np.float32(3.0)):
        got, ref = _run_both(src, np.float32, mask_w=k, mask_h=k, weights=weights, domain=dom,
                             reduce=red, boundary=mode, const_val=0.5, fp_contract=False)
        if mode == capi.UNDEFINED:
            got, ref = _interior(got, 2), _interior(ref, 2)
        np.testing.assert_array_equal(got, ref)
    src8 = ob.random_image((40, 90), np.uint8, 6)
    if red != capi.PROD:
        got, ref = _run_both(src8, np.uint8, mask_w=3, mask_h=3, acc_float=False, reduce=red,
                             boundary=mode, const_val=3)
        if mode == capi.UNDEFINED:
            got, ref = _interior(got, 1), _interior(ref, 1)
        np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("mode", [capi.CLAMP, capi.MIRROR, capi.REPEAT, capi.CONSTANT])
def test_windows_pitch_and_offsets(gpu, mode):
    """Accessor windows with offsets (boundary handling at the window edge),
    iteration spaces with offsets, pitched rows, unaligned windows."""
    print(f"windows mode={mode}", flush=True)
    src = ob.random_image((61, 333), np.float32, 99)
    k = 5
    w = masks.gaussian_2d(k, 1.0)
    for sw, dw, pitch in (((3, 2, 320, 55), (1, 4, 320, 55), 0),
                          ((8, 0, 256, 61), (0, 0, 256, 61), 384),
                          ((0, 5, 333, 50), (0, 0, 333, 50), 340)):
        got, ref = _run_both(src, np.float32, mask_w=k, mask_h=k, weights=w, boundary=mode,
                             const_val=-1.0, src_window=sw, dst_window=dw, pitch=pitch)
        np.testing.assert_array_equal(got, ref)


def test_batch_planes(gpu):
    src = ob.random_image((3, 50, 270), np.float32, 4)
    got, ref = _run_both(src, np.float32, mask_w=9, mask_h=9, weights=masks.gaussian_2d(9, 2.0),
                         boundary=capi.MIRROR)
    np.testing.assert_array_equal(got, ref)


def test_large_f32_5x5_clamp_matches_oracle(gpu):
    """The benchmark operator at its full size (4096x4096 float, 5x5 CLAMP)."""
    src = ob.random_image((4096, 4096), np.float32, 2024)
    w = masks.gaussian_2d(5, 1.0)
    got, ref = _run_both(src, np.float32, mask_w=5, mask_h=5, weights=w, boundary=capi.CLAMP)
    np.testing.assert_array_equal(got, ref)


def test_tiny_images_all_modes(gpu):
    """Windows narrower than the mask radius (pyramid tails)."""
    for (h, w) in ((1, 1), (2, 3), (3, 2), (4, 4), (7, 1)):
        src = ob.random_image((h, w), np.float32, h * 10 + w)
        for mode in (capi.CLAMP, capi.CONSTANT, capi.REPEAT):
            got, ref = _run_both(src, np.float32, mask_w=5, mask_h=5,
                                 weights=masks.gaussian_2d(5, 1.0), boundary=mode)
            np.testing.assert_array_equal(got, ref)


def test_invalid_arguments_are_reported(gpu):
    src = empty_gpu((16, 16), np.float32)
    dst = empty_gpu((16, 16), np.float32)
    with pytest.raises(capi.HipaccError) as e:
        op.local(src, dst, 4, 3, weights=np.ones(12, np.float32))
    assert e.value.status == 1
    with pytest.raises(capi.HipaccError):
        op.local(src, dst, 3, 3, weights=np.ones(9, np.float32), src_window=(10, 0, 10, 16))
