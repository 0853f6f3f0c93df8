"""Every kernel variant the autotuner can pick (stream with several segment
heights, tile with several rows per wave, ring with several block segment
heights, generic) produces the oracle's
result bit for bit, so the tuned choice never changes a result."""
import numpy as np
import pytest
import torch

from hipacc_amd import capi, masks, operators as op

import oracle_binding as ob
from gpu_util import empty_gpu, to_gpu, to_host
from weights import skewed_1d

pytestmark = pytest.mark.gpu

VARIANTS = [(capi.VAR_STREAM, 0), (capi.VAR_STREAM, 8), (capi.VAR_STREAM, 13),
            (capi.VAR_STREAM, 64), (capi.VAR_WIDE, 0), (capi.VAR_WIDE, 7), (capi.VAR_WIDE, 32),
            (capi.VAR_TILE, 2), (capi.VAR_TILE, 4), (capi.VAR_TILE, 8), (capi.VAR_GENERIC, 0),
            (capi.VAR_RING, 0), (capi.VAR_RING, 5), (capi.VAR_RING, 16), (capi.VAR_RING, 37)]
MODES = [capi.UNDEFINED, capi.CLAMP, capi.REPEAT, capi.MIRROR, capi.CONSTANT]


def _launch_all(call, check):
    ran = 0
    for v, p in VARIANTS:
        try:
            call.with_variant(v, p).launch()
        except capi.HipaccError as e:
            assert e.status == 2, e  # no kernel of this variant for the shape
            continue
        ran += 1
        check(v, p)
    assert ran >= 2


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("k", [3, 5, 7, 9, 13])
@pytest.mark.parametrize("shape,pitch", [((61, 500), 0), ((33, 1030), 1040), ((130, 97), 0)])
def test_f32_local_variants(gpu, mode, k, shape, pitch):
    print(f"variants f32 k={k} mode={mode} shape={shape}", flush=True)
    src = ob.random_image(shape, np.float32, k * 3 + shape[1])
    w = masks.gaussian_2d(k, k / 4.0)
    ref = np.zeros_like(src)
    ob.local(src, ref, k, k, weights=w, boundary=mode, const_val=0.7, fp_contract=True)
    dsrc = to_gpu(src, pitch)
    dst = empty_gpu(shape, np.float32, pitch)
    call = op.prepare_local(dsrc, dst, k, k, weights=w, boundary=mode, const_val=0.7)
    r = k // 2

    def check(v, p):
        got = to_host(dst, np.float32)
        g, e = (got[r:-r, r:-r], ref[r:-r, r:-r]) if mode == capi.UNDEFINED else (got, ref)
        np.testing.assert_array_equal(g, e, err_msg=f"variant {v}/{p}")
        dst.fill_(-1.0)
    _launch_all(call, check)


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("k", [3, 5, 7, 9])
def test_u8_local_variants(gpu, mode, k):
    src = ob.random_image((45, 777), np.uint8, k)
    w = masks.gaussian_2d(k, k / 4.0)
    ref = np.zeros_like(src)
    ob.local(src, ref, k, k, weights=w, boundary=mode, const_val=99, cvt=capi.CVT_ROUND,
             fp_contract=False)
    dst = empty_gpu(src.shape, np.uint8)
    call = op.prepare_local(to_gpu(src), dst, k, k, weights=w, boundary=mode, const_val=99,
                            cvt=capi.CVT_ROUND, fp_contract=False)
    r = k // 2

    def check(v, p):
        got = to_host(dst, np.uint8)
        g, e = (got[r:-r, r:-r], ref[r:-r, r:-r]) if mode == capi.UNDEFINED else (got, ref)
        np.testing.assert_array_equal(g, e, err_msg=f"variant {v}/{p}")
    _launch_all(call, check)


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("out", [np.int16, np.uint32, np.int32])
def test_int_sobel_variants(gpu, mode, out):
    src = ob.random_image((50, 600), np.uint8, 17)
    for m in (masks.SOBEL_X, masks.SOBEL_Y):
        dom = (m != 0).astype(np.uint8)
        ref = np.zeros(src.shape, out)
        ob.local(src, ref, 3, 3, weights=m, domain=dom, acc_float=False, boundary=mode,
                 const_val=5)
        dst = empty_gpu(src.shape, out)
        call = op.prepare_local(to_gpu(src), dst, 3, 3, weights=m, domain=dom, acc_float=False,
                                boundary=mode, const_val=5, dst_dtype=ob.CODES[np.dtype(out)])

        def check(v, p):
            got = to_host(dst, out)
            g, e = (got[1:-1, 1:-1], ref[1:-1, 1:-1]) if mode == capi.UNDEFINED else (got, ref)
            np.testing.assert_array_equal(g, e, err_msg=f"variant {v}/{p}")
        _launch_all(call, check)


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("k", [3, 5, 7, 9, 13])
@pytest.mark.parametrize("dt", [np.float32, np.uint8])
def test_separable_variants(gpu, mode, k, dt):
    src = ob.random_image((70, 1100), dt, 31 + k)
    g1 = masks.gaussian_1d(k, k / 4.0)
    c1 = skewed_1d(k, 60 + k)  # column taps differ from the row taps
    cvt = capi.CVT_ROUND if dt == np.uint8 else capi.CVT_TRUNC
    ref = np.zeros_like(src)
    ob.separable(src, ref, g1, c1, boundary=mode, const_val=3, cvt=cvt)
    dst = empty_gpu(src.shape, dt)
    call = op.prepare_separable(to_gpu(src), dst, g1, c1, boundary=mode, const_val=3, cvt=cvt)
    r = k // 2

    def check(v, p):
        got = to_host(dst, dt)
        g, e = (got[r:-r, r:-r], ref[r:-r, r:-r]) if mode == capi.UNDEFINED else (got, ref)
        np.testing.assert_array_equal(g, e, err_msg=f"variant {v}/{p}")
    _launch_all(call, check)


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("k", [3, 5, 7])
@pytest.mark.parametrize("out", [np.uint32, np.int16, np.uint8])
def test_box_variants(gpu, mode, k, out):
    src = ob.random_image((66, 900), np.uint8, 5 * k)
    ones = np.ones(k, np.int32)
    ref = np.zeros(src.shape, out)
    ob.local(src, ref, k, k, acc_float=False, boundary=mode, post_div=float(k * k)
             if out == np.uint8 else 0.0)
    dst = empty_gpu(src.shape, out)
    call = op.prepare_separable(to_gpu(src), dst, ones, ones, acc_float=False, boundary=mode,
                                dst_dtype=ob.CODES[np.dtype(out)],
                                post_div=float(k * k) if out == np.uint8 else 0.0)
    r = k // 2

    def check(v, p):
        got = to_host(dst, out)
        g, e = (got[r:-r, r:-r], ref[r:-r, r:-r]) if mode == capi.UNDEFINED else (got, ref)
        np.testing.assert_array_equal(g, e, err_msg=f"variant {v}/{p}")
    _launch_all(call, check)


def test_device_weights_variants(gpu):
    """Learnable masks: weights read from device memory."""
    src = ob.random_image((2, 64, 300), np.float32, 3)
    w = np.random.default_rng(1).standard_normal((9, 9)).astype(np.float32) * np.float32(0.1)
    for mode in (capi.CLAMP, capi.MIRROR):
        ref = np.zeros_like(src)
        ob.local(src, ref, 9, 9, weights=w, boundary=mode)
        dst = empty_gpu(src.shape, np.float32)
        wd = torch.from_numpy(w.ravel()).cuda()
        call = op.prepare_local(to_gpu(src), dst, 9, 9, weights_dev=wd, boundary=mode)

        def check(v, p):
            np.testing.assert_array_equal(to_host(dst, np.float32), ref, err_msg=f"{v}/{p}")
        _launch_all(call, check)


def test_tuner_records_a_winner_and_auto_uses_it(gpu, tmp_path):
    src = ob.random_image((512, 1024), np.float32, 9)
    w = masks.gaussian_2d(5, 1.0)
    dst = empty_gpu(src.shape, np.float32)
    call = op.prepare_local(to_gpu(src), dst, 5, 5, weights=w, boundary=capi.CLAMP)
    e = call.tune(iters=3)
    assert e.variant in capi.VARIANT_NAMES and e.ms > 0
    assert e.size_class == 19   # 512 * 1024 = 2^19 pixels
    # a large image (64 MiB per buffer set: the tuner rotates over 8 sets,
    # 512 MiB) gets its own entry in its own size class
    big = ob.random_image((2048, 4096), np.float32, 10)
    dbig = empty_gpu(big.shape, np.float32)
    cbig = op.prepare_local(to_gpu(big), dbig, 5, 5, weights=w, boundary=capi.CLAMP)
    eb = cbig.tune(iters=8)
    assert eb.size_class == 23 and eb.variant in capi.VARIANT_NAMES and eb.ms > 0
    path = str(tmp_path / "tune.txt").encode()
    capi.check(capi.lib().hipacc_hip_tune_save(path), "tune_save")
    lines = [ln.split() for ln in open(path.decode()) if not ln.startswith("#")]
    classes = {int(ln[9]) for ln in lines if ln[0] == "0" and ln[1] == "5" and ln[6] == "1"}
    assert {19, 23} <= classes, lines
    capi.check(capi.lib().hipacc_hip_tune_load(path), "tune_load")
    call.launch()
    cbig.launch()
    ref = np.zeros_like(src)
    ob.local(src, ref, 5, 5, weights=w, boundary=capi.CLAMP)
    np.testing.assert_array_equal(to_host(dst, np.float32), ref)
    refb = np.zeros_like(big)
    ob.local(big, refb, 5, 5, weights=w, boundary=capi.CLAMP)
    np.testing.assert_array_equal(to_host(dbig, np.float32), refb)
