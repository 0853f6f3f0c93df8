"""GPU parity of separable, point, reduction, histogram and bilateral
operators against the C oracle and the reference DSL golden vectors."""
import math

import numpy as np
import pytest
import torch

from hipacc_amd import capi, masks, operators as op

import golden_cases as gc
import oracle_binding as ob
from gpu_util import empty_gpu, to_gpu, to_host
from weights import skewed_1d

pytestmark = pytest.mark.gpu

MODES = [capi.CLAMP, capi.REPEAT, capi.MIRROR, capi.CONSTANT]


# ---- separable -----------------------------------------------------------
@pytest.mark.parametrize("mode", MODES + [capi.UNDEFINED])
@pytest.mark.parametrize("kx,ky", [(3, 3), (5, 5), (7, 7), (9, 9), (13, 13), (13, 7), (3, 9)])
def test_separable_f32_vs_two_pass_oracle(gpu, mode, kx, ky):
    """Row and column taps differ (a Gaussian row, a signed random column),
    so a swapped or transposed pass fails; kx != ky takes the shapes the
    square-mask kernels refuse."""
    print(f"separable f32 kx={kx} ky={ky} mode={mode}", flush=True)
    src = ob.random_image((77, 530), np.float32, 40 + kx + ky)
    g = masks.gaussian_1d(kx, kx / 4.0)
    c = skewed_1d(ky, 50 + ky, signed=True)
    for contract in (False, True):
        ref = np.zeros_like(src)
        ob.separable(src, ref, g, c, boundary=mode, const_val=0.3, fp_contract=contract)
        dst = empty_gpu(src.shape, np.float32)
        op.separable(to_gpu(src), dst, g, c, boundary=mode, const_val=0.3, fp_contract=contract)
        got = to_host(dst, np.float32)
        if mode == capi.UNDEFINED:
            rx, ry = kx // 2, ky // 2
            got, ref = got[ry:-ry, rx:-rx], ref[ry:-ry, rx:-rx]
        np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("mode", MODES)
def test_separable_u8_gaussian13(gpu, mode):
    src = ob.random_image((64, 2100), np.uint8, 77)
    g = masks.gaussian_1d(13, 3.0)
    c = skewed_1d(13, 78)  # a different column: pins the pass orientation
    ref = np.zeros_like(src)
    ob.separable(src, ref, g, c, boundary=mode, cvt=capi.CVT_ROUND, const_val=9)
    dst = empty_gpu(src.shape, np.uint8)
    op.separable(to_gpu(src), dst, g, c, boundary=mode, cvt=capi.CVT_ROUND, const_val=9)
    np.testing.assert_array_equal(to_host(dst, np.uint8), ref)


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("k", [3, 5, 7])
def test_box_sum_separable_equals_2d_reference(gpu, mode, k):
    """uint8 -> uint32 box sum: the separable integer path against the 2D
    Domain reduce of the oracle (exact: integer sums are associative).  With
    CONSTANT the two agree for the constant 0; for another constant the
    two-pass program differs by definition (its intermediate image's rows
    outside the window read the constant itself) and is checked against the
    two-pass oracle instead."""
    print(f"box k={k} mode={mode}", flush=True)
    src = ob.random_image((70, 1300), np.uint8, 5 + k)
    ones = np.ones(k, np.int32)
    for cval in (0, 250):
        ref = np.zeros(src.shape, np.uint32)
        if mode == capi.CONSTANT and cval:
            ob.separable(src, ref, ones, ones, acc_float=False, boundary=mode, const_val=cval)
        else:
            ob.local(src, ref, k, k, acc_float=False, boundary=mode, const_val=cval)
        dst = empty_gpu(src.shape, np.uint32)
        op.separable(to_gpu(src), dst, ones, ones, acc_float=False, boundary=mode, const_val=cval,
                     dst_dtype=capi.U32)
        np.testing.assert_array_equal(to_host(dst, np.uint32), ref)
        # the 2D entry point with an all-ones integer domain
        dst2 = empty_gpu(src.shape, np.uint32)
        op.local(to_gpu(src), dst2, k, k, acc_float=False, boundary=mode, const_val=cval,
                 dst_dtype=capi.U32)
        ref2 = np.zeros(src.shape, np.uint32)
        ob.local(src, ref2, k, k, acc_float=False, boundary=mode, const_val=cval)
        np.testing.assert_array_equal(to_host(dst2, np.uint32), ref2)


def test_box_7x7_golden(gpu):
    for c, e in gc.load():
        if c["op"] != "box_u8":
            continue
        src = gc.input_for(c)
        dst = empty_gpu(src.shape, np.uint32)
        ones = np.ones(7, np.int32)
        op.separable(to_gpu(src), dst, ones, ones, acc_float=False, boundary=c["boundary"],
                     dst_dtype=capi.U32)
        np.testing.assert_array_equal(to_host(dst, np.uint32), e)


# ---- point -----------------------------------------------------------------
@pytest.mark.parametrize("width", [203, 256])
@pytest.mark.parametrize("opcode", range(9))
@pytest.mark.parametrize("dt,odt", [(np.float32, np.float32), (np.uint8, np.uint8),
                                    (np.int16, np.uint8), (np.int8, np.int8),
                                    (np.float32, np.uint8)])
def test_point_vs_oracle(gpu, opcode, dt, odt, width):
    a = ob.random_image((37, width), dt, 1 + opcode)
    b = ob.random_image((37, width), dt, 100 + opcode)
    alpha = 4.0 if opcode == capi.POINT_MAGNITUDE else 1.7
    for cvt in (capi.CVT_TRUNC, capi.CVT_ROUND):
        ref = np.zeros(a.shape, odt)
        ob.point(opcode, a, ref, b, alpha=alpha, beta=0.25, cvt=cvt)
        dst = empty_gpu(a.shape, odt)
        op.point(opcode, to_gpu(a), dst, to_gpu(b), alpha=alpha, beta=0.25, cvt=cvt)
        np.testing.assert_array_equal(to_host(dst, odt), ref)


def test_point_windows_and_golden(gpu):
    for c, e in gc.load():
        if c["op"] != "sub_i8":
            continue
        a, b = gc.input_for(c)
        dst = empty_gpu(a.shape, np.int8)
        op.point(capi.POINT_SUB, to_gpu(a), dst, to_gpu(b))
        np.testing.assert_array_equal(to_host(dst, np.int8), e)
    a = ob.random_image((40, 100), np.float32, 3)
    ref = np.zeros((40, 100), np.float32)
    wins = ((5, 3, 60, 30), None, (20, 7, 60, 30))
    ob.point(capi.POINT_AFFINE, a, ref, alpha=2.0, beta=1.0, windows=wins)
    dst = empty_gpu((40, 100), np.float32)
    op.point(capi.POINT_AFFINE, to_gpu(a, pitch=128), dst, alpha=2.0, beta=1.0, windows=wins)
    np.testing.assert_array_equal(to_host(dst, np.float32), ref)


# ---- reductions ------------------------------------------------------------
def _gpu_reduce(src, mode, acc_dtype=None, window=None):
    res = torch.zeros(1, dtype=torch.float32 if src.dtype == np.float32 else torch.int32,
                      device="cuda")
    op.reduce(to_gpu(src), mode, res, acc_dtype=acc_dtype, window=window)
    torch.cuda.synchronize()
    return res.cpu().numpy()


@pytest.mark.parametrize("shape", [(1, 1), (3, 7), (29, 37), (300, 4099), (4096, 4096)])
def test_reduce_f32(gpu, shape):
    src = ob.random_image(shape, np.float32, shape[0] * 7 + shape[1])
    for mode in (capi.MIN, capi.MAX):
        got = _gpu_reduce(src, mode)[0]
        assert got == ob.reduce(src, mode)
    got = float(_gpu_reduce(src, capi.SUM)[0])
    exact = ob.sum_f64(src)
    # fp32 tree sum vs the FP64 sum: relative error bound ~ log2(n) * 2^-24
    tol = max(1e-6, 2 * math.log2(max(2, src.size)) * 2.0 ** -24)
    assert abs(got - exact) <= tol * max(1.0, abs(exact)), (got, exact)


@pytest.mark.parametrize("shape", [(1, 5), (29, 37), (1000, 3000)])
def test_reduce_u8(gpu, shape):
    src = ob.random_image(shape, np.uint8, 9 + shape[1])
    for mode in (capi.MIN, capi.MAX, capi.SUM):
        got = _gpu_reduce(src, mode, acc_dtype=capi.U32)[0]
        assert np.uint32(np.int32(got).view(np.uint32)) == ob.reduce(src, mode, acc_dtype=capi.U32)


def test_reduce_window_and_golden(gpu):
    src = ob.random_image((100, 200), np.float32, 8)
    win = (13, 7, 101, 55)
    assert _gpu_reduce(src, capi.MAX, window=win)[0] == ob.reduce(src, capi.MAX, window=win)
    for c, e in gc.load():
        if c["op"] == "red_f32" and c["args"][0] in ("1", "2"):
            mode = capi.MIN if c["args"][0] == "1" else capi.MAX
            assert _gpu_reduce(gc.input_for(c), mode)[0] == e[0]
        if c["op"] == "red_u8":
            mode = capi.MIN if c["args"][0] == "1" else capi.MAX
            assert _gpu_reduce(gc.input_for(c), mode, acc_dtype=capi.U32)[0] == e[0]


# ---- histogram ---------------------------------------------------------------
@pytest.mark.parametrize("shape", [(1, 1), (29, 37), (513, 1027), (4096, 4096)])
@pytest.mark.parametrize("nbins", [256, 64, 1000, 4096, 1])
def test_histogram_u8(gpu, shape, nbins):
    src = ob.random_image(shape, np.uint8, shape[1] + nbins)
    bins = torch.zeros(nbins, dtype=torch.int32, device="cuda")
    op.histogram(to_gpu(src), nbins, bins)
    got = to_host(bins, np.uint32)
    ref = ob.histogram(src, nbins)
    np.testing.assert_array_equal(got, ref)
    assert int(got.sum()) == src.size
    if nbins == 256:
        np.testing.assert_array_equal(got, np.bincount(src.ravel(), minlength=256))


def test_histogram_f32_and_golden(gpu):
    src = ob.random_image((300, 500), np.float32, 5) * np.float32(255.0)
    bins = torch.zeros(256, dtype=torch.int32, device="cuda")
    op.histogram(to_gpu(src), 256, bins, scale=256.0 / 255.0)
    np.testing.assert_array_equal(to_host(bins, np.uint32), ob.histogram(src, 256, 256.0 / 255.0))
    for c, e in gc.load():
        if c["op"] == "hist_u8":
            nb = int(c["args"][0])
            bins = torch.zeros(nb, dtype=torch.int32, device="cuda")
            op.histogram(to_gpu(gc.input_for(c)), nb, bins)
            np.testing.assert_array_equal(to_host(bins, np.uint32), e)


def test_histogram_constant_image(gpu):
    """All pixels in one bin: worst-case LDS atomic contention."""
    src = np.full((257, 1031), 200, np.uint8)
    bins = torch.zeros(256, dtype=torch.int32, device="cuda")
    op.histogram(to_gpu(src), 256, bins)
    got = to_host(bins, np.uint32)
    assert got[200] == src.size and got.sum() == src.size


# ---- bilateral -----------------------------------------------------------------
BILAT_REL = 1e-5


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("k,sd,sr", [(13, 3.0, 0.1), (7, 2.0, 0.2), (3, 1.0, 0.5)])
def test_bilateral_f32(gpu, mode, k, sd, sr):
    print(f"bilateral k={k} mode={mode}", flush=True)
    src = ob.random_image((53, 290), np.float32, 3 + k)
    ref = np.zeros_like(src)
    ob.bilateral(src, ref, k, sd, sr, boundary=mode, const_val=0.5)
    f64 = np.zeros(src.shape, np.float64)
    ob.bilateral(src, np.zeros_like(src), k, sd, sr, boundary=mode, const_val=0.5, f64_out=f64)
    dst = empty_gpu(src.shape, np.float32)
    op.bilateral(to_gpu(src), dst, k, sd, sr, boundary=mode, const_val=0.5)
    got = to_host(dst, np.float32)
    rel = np.abs(got.astype(np.float64) - ref) / np.maximum(np.abs(ref), 1e-6)
    assert rel.max() <= BILAT_REL, rel.max()
    err64 = np.abs(got - f64).max()
    ref64 = np.abs(ref - f64).max()
    print(f"bilateral k={k}: max|gpu-f64|={err64:.3e} max|cpu-f32 - f64|={ref64:.3e}")


def test_bilateral_golden(gpu):
    for c, e in gc.load():
        if c["op"] != "bilat_f32":
            continue
        k, sd, sr = int(c["args"][0]), float(c["args"][1]), float(c["args"][2])
        src = gc.input_for(c)
        dst = empty_gpu(src.shape, np.float32)
        op.bilateral(to_gpu(src), dst, k, sd, sr, boundary=c["boundary"])
        got = to_host(dst, np.float32)
        rel = np.abs(got.astype(np.float64) - e) / np.maximum(np.abs(e), 1e-6)
        assert rel.max() <= BILAT_REL, (gc.case_id(c), rel.max())


def test_bilateral_u8(gpu):
    src = ob.random_image((40, 300), np.uint8, 12)
    ref = np.zeros_like(src)
    ob.bilateral(src, ref, 5, 2.0, 20.0, boundary=capi.CLAMP, cvt=capi.CVT_ROUND)
    dst = empty_gpu(src.shape, np.uint8)
    op.bilateral(to_gpu(src), dst, 5, 2.0, 20.0, boundary=capi.CLAMP, cvt=capi.CVT_ROUND)
    got = to_host(dst, np.uint8).astype(int)
    # the float quotient can sit within rounding of a .5 boundary: +-1 LSB,
    # the reference samples' own integer tolerance
    assert np.abs(got - ref).max() <= 1
    assert (got != ref).mean() < 0.01


def test_fill_random_matches_host_generator(gpu):
    for dt in (np.uint8, np.float32, np.int16):
        t = empty_gpu((37, 101), dt, pitch=128)
        op.fill_random(t, 1234)
        np.testing.assert_array_equal(to_host(t, dt), ob.random_image((37, 101), dt, 1234))
