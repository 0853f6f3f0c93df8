"""Operators on 4-wide vector pixels (uchar4 / float4 in; uchar4, int4,
float4 out), the RGBA samples' operators: against the reference DSL's own
output (golden vectors) and the oracle, which applies the scalar operator
to each channel (the DSL's vector arithmetic is componentwise).  Point
operators on vector pixels run as elementwise operators on the channels."""
import numpy as np
import pytest
import torch

import golden_cases as gc
import oracle_binding as ob
from gpu_util import to_host
from hipacc_amd import capi, masks, operators as op

pytestmark = pytest.mark.gpu

VEC_CASES = [(c, e) for c, e in gc.load()
             if c["op"] in ("conv_rgba", "conv_f32x4", "box_rgba", "erode_rgba", "sobel_rgba")]
TORCH = {np.dtype(np.uint8): torch.uint8, np.dtype(np.int32): torch.int32,
         np.dtype(np.float32): torch.float32, np.dtype(np.int16): torch.int16}


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.mark.parametrize("case,expected", VEC_CASES, ids=[gc.case_id(c) for c, _ in VEC_CASES])
def test_vec4_golden(gpu, case, expected):
    src = gc.input_for(case)
    spec = gc.local_spec(case)
    out_dtype = spec.pop("out_dtype")
    for key in ("fp_contract",):
        spec.setdefault(key, False)
    dst = torch.zeros(expected.shape, dtype=TORCH[np.dtype(out_dtype)], device="cuda")
    op.local(_dev(src), dst, **spec)
    np.testing.assert_array_equal(to_host(dst, out_dtype), expected)


@pytest.mark.parametrize("mode", [capi.CLAMP, capi.MIRROR, capi.REPEAT, capi.CONSTANT,
                                  capi.UNDEFINED])
@pytest.mark.parametrize("k", [3, 5, 9, 15])
def test_vec4_conv_vs_oracle(gpu, mode, k):
    """uchar4 -> uchar4 rounded Gaussian and float4 -> float4 weighted sums
    with contraction on and off, a cropped accessor window into an offset
    iteration space, odd sizes."""
    print(f"vec4 conv k={k} mode={mode}", flush=True)
    h, w = 75, 133
    u8 = ob.random_image((h, 4 * w), np.uint8, 500 + k).reshape(h, w, 4)
    f32 = ob.random_image((h, 4 * w), np.float32, 600 + k).reshape(h, w, 4)
    g = masks.gaussian_2d(k, k / 4.0)
    win, owin = (7, 5, 100, 60), (3, 2, 100, 60)
    for contract in (False, True):
        ref = np.zeros((70, 110, 4), np.uint8)
        ob.local(u8, ref, k, k, weights=g, boundary=mode, const_val=17, cvt=capi.CVT_ROUND,
                 fp_contract=contract, src_window=win, dst_window=owin, src_dtype=capi.U8X4,
                 dst_dtype=capi.U8X4)
        dst = torch.zeros((70, 110, 4), dtype=torch.uint8, device="cuda")
        op.local(_dev(u8), dst, k, k, weights=g, boundary=mode, const_val=17,
                 cvt=capi.CVT_ROUND, fp_contract=contract, src_window=win, dst_window=owin,
                 src_dtype=capi.U8X4, dst_dtype=capi.U8X4)
        got = to_host(dst, np.uint8)
        if mode == capi.UNDEFINED:
            r = k // 2
            got, ref = got[2 + r:62 - r, 3 + r:103 - r], ref[2 + r:62 - r, 3 + r:103 - r]
        np.testing.assert_array_equal(got, ref)
        reff = np.zeros_like(f32)
        ob.local(f32, reff, k, k, weights=g, boundary=mode, const_val=0.25, fp_contract=contract,
                 src_dtype=capi.F32X4, dst_dtype=capi.F32X4)
        dstf = torch.zeros_like(_dev(f32))
        op.local(_dev(f32), dstf, k, k, weights=g, boundary=mode, const_val=0.25,
                 fp_contract=contract, src_dtype=capi.F32X4, dst_dtype=capi.F32X4)
        gotf = to_host(dstf, np.float32)
        if mode == capi.UNDEFINED:
            r = k // 2
            gotf, reff = gotf[r:-r, r:-r], reff[r:-r, r:-r]
        np.testing.assert_array_equal(gotf, reff)


@pytest.mark.parametrize("mode", [capi.CLAMP, capi.MIRROR, capi.CONSTANT])
def test_vec4_integer_and_minmax_vs_oracle(gpu, mode):
    """Box sums into int4 and uchar4 (/ k*k), Laplace-style signed integer
    masks into int4, and 7x5 erode / dilate on uchar4."""
    h, w = 64, 97
    u8 = ob.random_image((h, 4 * w), np.uint8, 700).reshape(h, w, 4)
    lap = np.array([[0, 1, 0], [1, -4, 1], [0, 1, 0]], np.int32)
    cases = [
        dict(mask_w=5, mask_h=5, acc_float=False, out=(np.int32, capi.I32X4)),
        dict(mask_w=5, mask_h=5, acc_float=False, post_div=25.0, out=(np.uint8, capi.U8X4)),
        dict(mask_w=3, mask_h=3, weights=lap, acc_float=False, out=(np.int32, capi.I32X4)),
        dict(mask_w=7, mask_h=5, acc_float=False, reduce=capi.MIN, out=(np.uint8, capi.U8X4)),
        dict(mask_w=7, mask_h=5, acc_float=False, reduce=capi.MAX, out=(np.uint8, capi.U8X4)),
    ]
    for c in cases:
        npd, code = c.pop("out")
        ref = np.zeros((h, w, 4), npd)
        ob.local(u8, ref, boundary=mode, const_val=9, src_dtype=capi.U8X4, dst_dtype=code, **c)
        dst = torch.zeros((h, w, 4), dtype=TORCH[np.dtype(npd)], device="cuda")
        op.local(_dev(u8), dst, boundary=mode, const_val=9, src_dtype=capi.U8X4, dst_dtype=code,
                 **c)
        np.testing.assert_array_equal(to_host(dst, npd), ref, err_msg=str(c))


def test_vec4_point_ops(gpu):
    """Point operators on uchar4 / float4: the Sobel_RGBA combine (magnitude
    with a norm) and an absolute difference, channel by channel."""
    h, w = 40, 77
    a = ob.random_image((h, 4 * w), np.int32, 800).reshape(h, w, 4) % 2000 - 1000
    b = ob.random_image((h, 4 * w), np.int32, 801).reshape(h, w, 4) % 2000 - 1000
    ref = np.zeros((h, 4 * w), np.uint8)
    ob.point(capi.POINT_MAGNITUDE, a.reshape(h, 4 * w).astype(np.int32), ref,
             b.reshape(h, 4 * w).astype(np.int32), alpha=4.0)
    dst = torch.zeros((h, w, 4), dtype=torch.uint8, device="cuda")
    op.point(capi.POINT_MAGNITUDE, _dev(a.astype(np.int32)), dst, _dev(b.astype(np.int32)),
             alpha=4.0, src_dtype=capi.I32X4, dst_dtype=capi.U8X4)
    np.testing.assert_array_equal(to_host(dst, np.uint8).reshape(h, 4 * w), ref)
    # mixing vector and scalar pixels is refused
    with pytest.raises(capi.HipaccError):
        op.point(capi.POINT_COPY, _dev(a.astype(np.int32)), dst, src_dtype=capi.I32X4)


def test_vec4_refusals(gpu):
    """Vector and scalar pixels do not mix; separable masks and reductions
    take scalar pixels only."""
    u8 = _dev(np.zeros((8, 8, 4), np.uint8))
    with pytest.raises(capi.HipaccError):
        op.local(u8, torch.zeros((8, 32), dtype=torch.uint8, device="cuda"), 3, 3,
                 src_dtype=capi.U8X4, acc_float=False)
    with pytest.raises(capi.HipaccError):
        op.separable(u8, torch.zeros_like(u8), np.ones(3, np.int32), np.ones(3, np.int32),
                     acc_float=False, src_dtype=capi.U8X4, dst_dtype=capi.U8X4)
