"""GPU parity of fused operator chains (hipacc_hip_fused) against the reference
DSL's unfused chains (golden vectors of the fusion samples) and against the
oracle run stage by stage.

Bar: bit-exact.  A fused launch must produce what the unfused chain of
hipacc_hip_point / hipacc_hip_local calls produces, intermediate pixel types
and boundary handling included.
"""
import ctypes

import numpy as np
import pytest
import torch

from hipacc_amd import capi, masks, operators as op

import golden_cases as gc
import oracle_binding as ob
from gpu_util import empty_gpu, to_gpu, to_host
from weights import skewed_1d

pytestmark = pytest.mark.gpu

GOLDEN = [(c, e) for c, e in gc.load() if c["op"] in ("fuse_u8", "fuse_f32")]
MODES = [capi.UNDEFINED, capi.CLAMP, capi.REPEAT, capi.MIRROR, capi.CONSTANT]


def _gpu_stages(stages):
    out = []
    for st in stages:
        st = dict(st)
        if isinstance(st.get("in1"), np.ndarray):
            st["in1"] = to_gpu(st["in1"])
        out.append(st)
    return out


def _fused(src, stages, *, src_window=None, pitch=0):
    shape = src.shape if src_window is None else src.shape[:-2] + (src_window[3], src_window[2])
    out_dt = ob.NP_DTYPES[stages[-1]["dtype"]]
    dst = empty_gpu(shape, out_dt, pitch=pitch, fill=0x5A)
    op.fused(to_gpu(src, pitch=pitch), dst, _gpu_stages(stages), src_window=src_window,
             dst_dtype=stages[-1]["dtype"])
    return to_host(dst, out_dt)


def _check(src, stages, **kw):
    ref = ob.chain(src, stages, src_window=kw.get("src_window"))
    got = _fused(src, stages, **kw)
    np.testing.assert_array_equal(got, ref)
    return got


@pytest.mark.parametrize("case,expected", GOLDEN, ids=[gc.case_id(c) for c, _ in GOLDEN])
def test_fused_matches_reference_dsl(gpu, case, expected):
    """The fusion samples' chains (P2P, P2L, L2P, L2L, Mixed) in one launch
    equal the reference's unfused DSL chain."""
    print(f"golden {gc.case_id(case)}", flush=True)
    got = _fused(gc.input_for(case), gc.fuse_spec(case))
    np.testing.assert_array_equal(got, expected)


def _gauss(k, sigma=1.0):
    return masks.gaussian_2d(k, sigma).reshape(-1)


def _local(dt, k, mode, **kw):
    kx, ky = (k, k) if isinstance(k, int) else k
    d = dict(kind="local", dtype=dt, mask_w=kx, mask_h=ky, boundary=mode)
    d.update(kw)
    return d


def _point(dt, op_=capi.POINT_AFFINE, **kw):
    d = dict(kind="point", dtype=dt, op=op_)
    d.update(kw)
    return d


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("shape", [(45, 97), (16, 64), (67, 130), (3, 5), (2, 40, 33)])
def test_fused_float_l2l(gpu, mode, shape):
    """f32 point -> 5x5 -> point -> 3x7 -> point, both locals with the same
    boundary mode, contraction on; sizes with partial tiles, an image
    smaller than the masks and a batch of two."""
    print(f"l2l f32 mode={mode} shape={shape}", flush=True)
    src = ob.random_image(shape, np.float32, 11)
    row, col = skewed_1d(3, 5), skewed_1d(7, 6)
    w2 = np.outer(col, row).reshape(-1).astype(np.float32)
    stages = [_point(capi.F32, alpha=1.5, beta=-0.25),
              _local(capi.F32, 5, mode, weights=_gauss(5), const_val=0.75),
              _point(capi.F32, alpha=0.5, beta=0.125),
              _local(capi.F32, (3, 7), mode, weights=w2, const_val=-2.0),
              _point(capi.F32, alpha=2.0, beta=1.0)]
    _check(src, stages)


@pytest.mark.parametrize("m1,m2", [(capi.CLAMP, capi.REPEAT), (capi.REPEAT, capi.MIRROR),
                                   (capi.CONSTANT, capi.CLAMP), (capi.MIRROR, capi.CONSTANT),
                                   (capi.UNDEFINED, capi.REPEAT)])
@pytest.mark.parametrize("shape", [(29, 37), (70, 131), (4, 6)])
def test_fused_mixed_modes_u8(gpu, m1, m2, shape):
    """The Mixed sample's shape on uchar images (P P L P P L P P) with two
    different boundary modes, fp_contract off and on."""
    print(f"mixed u8 {m1} {m2} {shape}", flush=True)
    src = ob.random_image(shape, np.uint8, 23)
    for fma in (False, True):
        p = _point(capi.U8, alpha=1.0, beta=3.0)
        stages = [p, p, _local(capi.U8, 5, m1, weights=gc.FUSE_MASK.reshape(-1),
                               cvt=capi.CVT_ROUND, fp_contract=fma, const_val=200),
                  p, p, _local(capi.U8, 7, m2, weights=_gauss(7, 1.5), cvt=capi.CVT_ROUND,
                               fp_contract=fma, const_val=17),
                  p, p]
        _check(src, stages)


@pytest.mark.parametrize("mode", MODES)
def test_fused_integer_types(gpu, mode):
    """u8 -> COPY to i16 -> Sobel-x (int acc) -> ABSDIFF with a second i16
    image -> ROUND/2 affine to u8 -> MIN over a 3x3 domain with holes."""
    print(f"int chain mode={mode}", flush=True)
    h, w = 53, 77
    src = ob.random_image((h, w), np.uint8, 5)
    other = ob.random_image((h, w), np.int16, 6) // 64
    sob = np.array([-1, 0, 1, -2, 0, 2, -1, 0, 1], np.int32)
    dom = np.array([1, 0, 1, 1, 1, 1, 0, 1, 1], np.uint8)
    stages = [_point(capi.I16, capi.POINT_COPY),
              _local(capi.I16, 3, mode, weights=sob, acc_float=False, const_val=-9),
              _point(capi.I16, capi.POINT_ABSDIFF, in1=other),
              _point(capi.U8, alpha=0.5, beta=0.0, cvt=capi.CVT_ROUND),
              _local(capi.U8, 3, mode, domain=dom, acc_float=False, reduce=capi.MIN,
                     const_val=250)]
    _check(src, stages)


@pytest.mark.parametrize("reduce", [capi.SUM, capi.MIN, capi.MAX, capi.PROD])
@pytest.mark.parametrize("k", [(1, 1), (15, 15), (3, 9)])
def test_fused_reduce_and_mask_shapes(gpu, reduce, k):
    """Unweighted reductions and weighted sums with post_div, masks from
    1x1 to 15x15, f32 -> i32 -> f32 types."""
    print(f"reduce={reduce} k={k}", flush=True)
    src = ob.random_image((41, 83), np.float32, 8)
    kx, ky = k
    dom = (np.arange(kx * ky) % 5 != 3).astype(np.uint8)
    dom[(kx * ky) // 2] = 1
    wts = None if reduce != capi.SUM else np.linspace(-1, 1, kx * ky).astype(np.float32)
    stages = [_point(capi.I32, alpha=100.0, beta=-50.0),
              _local(capi.F32, k, capi.MIRROR, weights=wts, domain=dom, reduce=reduce,
                     acc_float=True, post_div=3.0 if reduce == capi.SUM else 0.0),
              _point(capi.I16, capi.POINT_COPY)]
    if reduce == capi.PROD:
        stages[0] = _point(capi.I32, alpha=0.0, beta=1.0)    # products of ones
    _check(src, stages)


def test_fused_first_local_reads_window(gpu):
    """A chain starting with a local reads its input through the accessor
    window (UNDEFINED reads the pixels beyond the window, as
    hipacc_hip_local), writes into an offset window of a pitched output."""
    big = ob.random_image((60, 90), np.float32, 31)
    win = (7, 5, 61, 43)
    for mode in MODES:
        print(f"window mode={mode}", flush=True)
        stages = [_local(capi.F32, 7, mode, weights=_gauss(7, 2.0), const_val=0.3),
                  _point(capi.U8, alpha=255.0, cvt=capi.CVT_ROUND)]
        ref = ob.chain(big, stages, src_window=win)
        dst = empty_gpu((50, 80), np.uint8, pitch=96, fill=0x77)
        op.fused(to_gpu(big, pitch=128), dst, stages, src_window=win,
                 dst_window=(9, 4, 61, 43))
        got = to_host(dst, np.uint8)
        np.testing.assert_array_equal(got[4:47, 9:70], ref)
        outside = np.ones(got.shape, bool)
        outside[4:47, 9:70] = False
        assert (got[outside] == 0x77).all()


def test_fused_points_only_and_magnitude(gpu):
    """P2P chain with two-input ops (MAGNITUDE on u8 with a norm, MAX)."""
    h, w = 33, 65
    src = ob.random_image((h, w), np.uint8, 41)
    b = ob.random_image((h, w), np.uint8, 42)
    stages = [_point(capi.U8, capi.POINT_MAGNITUDE, alpha=2.0, in1=b),
              _point(capi.U8, capi.POINT_MAX, in1=b),
              _point(capi.F32, alpha=1.0 / 255.0, beta=0.0)]
    _check(src, stages)


def test_fused_equals_unfused_gpu_chain(gpu):
    """The fused launch against the same chain run as separate
    hipacc_hip_point / hipacc_hip_local launches on the GPU."""
    src = ob.random_image((2, 300, 257), np.float32, 51)
    stages = [_point(capi.F32, alpha=0.9, beta=0.05),
              _local(capi.F32, 5, capi.CLAMP, weights=_gauss(5)),
              _point(capi.F32, alpha=2.0, beta=-0.5),
              _local(capi.F32, 9, capi.MIRROR, weights=_gauss(9, 2.0)),
              _point(capi.F32, capi.POINT_COPY)]
    cur = to_gpu(src)
    for st in stages:
        st = dict(st)
        kind, _ = st.pop("kind"), st.pop("dtype")
        nxt = torch.empty_like(cur)
        if kind == "point":
            op.point(st.pop("op"), cur, nxt, alpha=st.pop("alpha", 1.0), beta=st.pop("beta", 0.0))
        else:
            op.local(cur, nxt, st.pop("mask_w"), st.pop("mask_h"), **st)
        cur = nxt
    got = _fused(src, stages)
    np.testing.assert_array_equal(got, to_host(cur, np.float32))


def _expect_error(src, dst, stages, **kw):
    with pytest.raises(capi.HipaccError):
        op.fused(src, dst, stages, **kw)


def test_fused_refusals(gpu):
    src = to_gpu(ob.random_image((20, 30), np.float32, 1))
    dst = empty_gpu((20, 30), np.float32)
    loc = _local(capi.F32, 3, capi.CLAMP, weights=_gauss(3))
    _expect_error(src, dst, [loc, loc, loc])                                  # three locals
    with pytest.raises(ValueError):
        op.fused(src, dst, [_point(capi.F32)] * 9)                             # nine stages
    d = op.prepare_fused(src, dst, [_point(capi.F32)]).desc
    d.num_stages = 9                                                          # past the table
    assert capi.lib().hipacc_hip_fused(ctypes.byref(d), None) == 1   # HIPACC_ERR_INVALID_ARG
    _expect_error(src, dst, [_point(capi.U8)])                                # out pixel type
    _expect_error(src, dst, [_point(capi.F32, capi.POINT_ADD,
                                    in1=to_gpu(ob.random_image((20, 30), np.uint8, 2)))])
    _expect_error(src, dst, [_point(capi.F32, capi.POINT_ADD)])               # no second input
    _expect_error(src, empty_gpu((20, 31), np.float32), [loc])                # sizes differ
    w = torch.zeros(9, device="cuda")
    _expect_error(src, dst, [_local(capi.F32, 3, capi.CLAMP, weights_dev=w)])  # device weights
    _expect_error(src, dst, [_local(capi.F32, 4, capi.CLAMP, weights=np.ones(16))])  # even mask
