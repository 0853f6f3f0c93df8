"""Pins the C oracle against what the mathematics fixes, independently of the
reference DSL fixtures:

* boundary modes against scipy.ndimage / numpy.pad closed forms
  (CLAMP='nearest'/'edge', REPEAT='wrap', MIRROR='reflect'/'symmetric' -- the
  edge-duplicating reflection, CONSTANT='constant'),
* correlation orientation (mask index [dy+r][dx+r] multiplies in(x+dx, y+dy)),
* the two float rounding orders (mul-then-add, and fmaf) by exact rational
  arithmetic on sampled pixels,
* the float separable program (row pass into a float32 intermediate, then
  column pass) by exact rational arithmetic and against FP64 correlate1d,
* integer kernels, separable == 2D outer product, reductions, histograms,
  bilateral and point operators by brute force.
A planted mistake (wrong reflection, transposed mask, dropped tap, FMA in the
no-contract path) fails at least one of these.
"""
import math
from fractions import Fraction

import numpy as np
import pytest
import scipy.ndimage as ndi

from hipacc_amd import capi, masks

import oracle_binding as ob
from weights import skewed_1d

SCIPY_MODE = {capi.CLAMP: "nearest", capi.REPEAT: "wrap", capi.MIRROR: "reflect",
              capi.CONSTANT: "constant"}
PAD_MODE = {capi.CLAMP: "edge", capi.REPEAT: "wrap", capi.MIRROR: "symmetric",
            capi.CONSTANT: "constant"}


@pytest.mark.parametrize("mode", list(SCIPY_MODE))
@pytest.mark.parametrize("k", [3, 5, 9])
def test_boundary_and_orientation_vs_scipy(mode, k):
    src = ob.random_image((23, 31), np.float32, 5 + k)
    rng = np.random.default_rng(k)
    w = rng.standard_normal((k, k)).astype(np.float32)  # asymmetric: pins orientation
    f64 = np.zeros(src.shape, np.float64)
    ob.local(src, np.zeros_like(src), k, k, weights=w, boundary=mode, const_val=0.375, f64_out=f64)
    ref = ndi.correlate(src.astype(np.float64), w.astype(np.float64), mode=SCIPY_MODE[mode],
                        cval=0.375)
    np.testing.assert_allclose(f64, ref, rtol=0, atol=1e-12)


@pytest.mark.parametrize("mode", list(SCIPY_MODE))
def test_boundary_applies_to_accessor_window(mode):
    """Boundary handling refers to the accessor window, not the image."""
    src = ob.random_image((30, 40), np.float32, 77)
    win = (5, 4, 25, 20)
    w = np.arange(25, dtype=np.float32).reshape(5, 5) / 300
    f64 = np.zeros((20, 25), np.float64)
    ob.local(src, np.zeros((20, 25), np.float32), 5, 5, weights=w, boundary=mode, const_val=-2.0,
             src_window=win, f64_out=f64)
    crop = src[4:24, 5:30].astype(np.float64)
    ref = ndi.correlate(crop, w.astype(np.float64), mode=SCIPY_MODE[mode], cval=-2.0)
    np.testing.assert_allclose(f64, ref, atol=1e-12)


def _round_f32(q: Fraction) -> np.float32:
    """Round an exact rational to the nearest float32, ties to even."""
    f = np.float32(float(q))
    best = None
    for c in (np.nextafter(f, np.float32(-np.inf)), f, np.nextafter(f, np.float32(np.inf))):
        d = abs(Fraction(float(c)) - q)
        if best is None or d < best[0] or (d == best[0] and (int(np.float32(c).view(np.uint32)) & 1) == 0):
            best = (d, np.float32(c))
    return best[1]


@pytest.mark.parametrize("contract", [False, True])
def test_float_rounding_order_exact(contract):
    src = ob.random_image((12, 14), np.float32, 3)
    w = masks.gaussian_2d(5, 1.0)
    dst = np.zeros_like(src)
    ob.local(src, dst, 5, 5, weights=w, boundary=capi.CLAMP, fp_contract=contract)
    pad = np.pad(src, 2, mode="edge")
    for (y, x) in ((0, 0), (3, 7), (11, 13), (6, 2), (9, 9)):
        acc = None
        for dy in range(5):
            for dx in range(5):
                wv, xv = Fraction(float(w[dy, dx])), Fraction(float(pad[y + dy, x + dx]))
                if acc is None:
                    acc = _round_f32(wv * xv)
                elif contract:
                    acc = _round_f32(wv * xv + Fraction(float(acc)))
                else:
                    p = _round_f32(wv * xv)
                    acc = _round_f32(Fraction(float(p)) + Fraction(float(acc)))
        assert dst[y, x] == acc, (y, x, dst[y, x], acc)


@pytest.mark.parametrize("mode", list(PAD_MODE))
def test_integer_kernels_brute_force(mode):
    src = ob.random_image((19, 27), np.uint8, 8)
    pad = np.pad(src.astype(np.int64), 3, mode=PAD_MODE[mode],
                 **({"constant_values": 11} if mode == capi.CONSTANT else {}))
    # 7x7 box sum -> uint32
    box = np.zeros(src.shape, np.uint32)
    ob.local(src, box, 7, 7, acc_float=False, boundary=mode, const_val=11)
    ref = sum(pad[dy:dy + 19, dx:dx + 27] for dy in range(7) for dx in range(7))
    np.testing.assert_array_equal(box, ref.astype(np.uint32))
    # Sobel-x with its Domain (zero taps skipped) -> int16
    sob = np.zeros(src.shape, np.int16)
    ob.local(src, sob, 3, 3, weights=masks.SOBEL_X, domain=(masks.SOBEL_X != 0).astype(np.uint8),
             acc_float=False, boundary=mode, const_val=11)
    p1 = pad[2:-2, 2:-2]
    ref = sum(int(masks.SOBEL_X[dy, dx]) * p1[dy:dy + 19, dx:dx + 27]
              for dy in range(3) for dx in range(3))
    np.testing.assert_array_equal(sob, ref.astype(np.int16))
    # 3x3 erode (MIN)
    mn = np.zeros(src.shape, np.uint8)
    ob.local(src, mn, 3, 3, acc_float=False, reduce=capi.MIN, boundary=mode, const_val=11)
    ref = np.min(np.stack([p1[dy:dy + 19, dx:dx + 27] for dy in range(3) for dx in range(3)]), 0)
    np.testing.assert_array_equal(mn, ref.astype(np.uint8))


@pytest.mark.parametrize("mode", list(PAD_MODE))
def test_separable_integer_equals_2d_outer_product(mode):
    src = ob.random_image((21, 33), np.uint8, 4)
    row = np.array([1, 2, 3, 2, 1], np.int32)
    col = np.array([-1, 0, 4, 0, -1], np.int32)
    sep = np.zeros(src.shape, np.int32)
    ob.separable(src, sep, row, col, acc_float=False, boundary=mode, const_val=0)
    full = np.zeros(src.shape, np.int32)
    ob.local(src, full, 5, 5, weights=np.outer(col, row), acc_float=False, boundary=mode,
             const_val=0)
    if mode != capi.CONSTANT:  # CONSTANT: the intermediate's own constant differs
        np.testing.assert_array_equal(sep, full)
    # the interior agrees in every mode
    np.testing.assert_array_equal(sep[2:-2, 2:-2], full[2:-2, 2:-2])


def test_reductions_and_histogram_brute_force():
    src = ob.random_image((50, 70), np.float32, 10)
    assert ob.reduce(src, capi.MIN) == src.min()
    assert ob.reduce(src, capi.MAX) == src.max()
    assert ob.sum_f64(src) == pytest.approx(math.fsum(src.ravel().astype(np.float64)), abs=1e-9)
    # sequential float sum (DSL order) == numpy cumulative float32 sum
    assert ob.reduce(src, capi.SUM, nthreads=1) == np.cumsum(src.ravel(), dtype=np.float32)[-1]
    u8 = ob.random_image((50, 70), np.uint8, 11)
    assert ob.reduce(u8, capi.SUM, acc_dtype=capi.U32, nthreads=4) == int(u8.sum())
    np.testing.assert_array_equal(ob.histogram(u8, 256), np.bincount(u8.ravel(), minlength=256))
    np.testing.assert_array_equal(ob.histogram(u8, 32),
                                  np.bincount(u8.ravel().astype(int) * 32 // 256, minlength=32))


def test_bilateral_f64_brute_force():
    src = ob.random_image((9, 11), np.float32, 12)
    k, sd, sr = 5, 1.5, 0.2
    f64 = np.zeros(src.shape, np.float64)
    ob.bilateral(src, np.zeros_like(src), k, sd, sr, boundary=capi.MIRROR, f64_out=f64)
    pad = np.pad(src.astype(np.float64), 2, mode="symmetric")
    sp = masks.bilateral_spatial(k, sd).astype(np.float64)
    for y in range(9):
        for x in range(11):
            win = pad[y:y + 5, x:x + 5]
            wgt = np.exp(-0.5 / (np.float64(np.float32(sr)) ** 2) * (win - src[y, x]) ** 2) * sp
            assert f64[y, x] == pytest.approx((wgt * win).sum() / wgt.sum(), rel=1e-12)
    # constant image: the filter is the identity
    c = np.full((8, 8), 0.3, np.float32)
    out = np.zeros_like(c)
    ob.bilateral(c, out, 5, 1.5, 0.2, boundary=capi.CLAMP)
    np.testing.assert_allclose(out, c, rtol=1e-6)


def test_point_ops_brute_force():
    a = ob.random_image((7, 9), np.int16, 1)
    b = ob.random_image((7, 9), np.int16, 2)
    out = np.zeros((7, 9), np.uint8)
    ob.point(capi.POINT_MAGNITUDE, a, out, b, alpha=4.0)
    qa, qb = np.trunc(a / 4).astype(np.int64), np.trunc(b / 4).astype(np.int64)
    ref = np.clip(np.sqrt((qa * qa + qb * qb).astype(np.float32)), 0, 255).astype(np.uint8)
    np.testing.assert_array_equal(out, ref)
    i8a = ob.random_image((7, 9), np.int8, 3)
    i8b = ob.random_image((7, 9), np.int8, 4)
    o8 = np.zeros((7, 9), np.int8)
    ob.point(capi.POINT_SUB, i8a, o8, i8b)
    np.testing.assert_array_equal(o8, (i8a.astype(np.int32) - i8b).astype(np.int8))


def test_output_conversion_semantics():
    """(uchar)(x + 0.5f) rounding and out-of-range wrap of the cvt modes."""
    vals = np.array([[0.49, 0.5, 254.5, 255.49, -1.2, 256.0, 300.7, -0.4]], np.float32)
    out = np.zeros(vals.shape, np.uint8)
    ob.point(capi.POINT_COPY, vals, out, cvt=capi.CVT_ROUND)
    np.testing.assert_array_equal(out[0], [0, 1, 255, 255, 0, 0, 45, 0])
    ob.point(capi.POINT_COPY, vals, out, cvt=capi.CVT_TRUNC)
    np.testing.assert_array_equal(out[0], [0, 0, 254, 255, 255, 0, 44, 0])


# ---- float separable: both passes pinned independently of the oracle ------
SEP_MODES = [capi.CLAMP, capi.REPEAT, capi.MIRROR, capi.CONSTANT]


def _index_map(n: int, r: int, mode: int) -> np.ndarray:
    """Source index of each padded position -r .. n+r-1 (numpy.pad semantics
    for the mode; -1 marks the CONSTANT value)."""
    if mode == capi.CONSTANT:
        return np.pad(np.arange(n), r, mode="constant", constant_values=-1)
    return np.pad(np.arange(n), r, mode=PAD_MODE[mode])


def _exact_sum(weights, values, contract):
    """Row-major accumulation of w*v as float32 code does it: the first
    product seeds the sum, then either one rounding per fused multiply-add
    (contract) or a rounded product followed by a rounded add."""
    acc = None
    for wv, xv in zip(weights, values):
        pw, px = Fraction(float(wv)), Fraction(float(xv))
        if acc is None:
            acc = _round_f32(pw * px)
        elif contract:
            acc = _round_f32(pw * px + Fraction(float(acc)))
        else:
            p = _round_f32(pw * px)
            acc = _round_f32(Fraction(float(p)) + Fraction(float(acc)))
    return acc


@pytest.mark.parametrize("mode", SEP_MODES)
@pytest.mark.parametrize("kx,ky", [(13, 7), (5, 9), (3, 3)])
@pytest.mark.parametrize("contract", [False, True])
def test_separable_float_rounding_order_exact(mode, kx, ky, contract):
    """Each output is the column pass over float32 intermediate pixels, each
    intermediate pixel the row pass over input pixels, both accumulated in
    the stated rounding order; intermediate rows outside the window read the
    CONSTANT value.  Row and column taps differ and kx != ky, so swapping
    the passes or the tap order fails."""
    H, W = 11, 17
    src = ob.random_image((H, W), np.float32, 100 + kx)
    row = skewed_1d(kx, 7 + kx, signed=True)
    col = skewed_1d(ky, 9 + ky, signed=True)
    cval = 0.375
    dst = np.zeros_like(src)
    ob.separable(src, dst, row, col, boundary=mode, const_val=cval, fp_contract=contract)
    rx, ry = kx // 2, ky // 2
    xm, ym = _index_map(W, rx, mode), _index_map(H, ry, mode)

    def inter(yy, x):  # intermediate pixel (row pass) at source row yy
        if yy < 0:
            return np.float32(cval)
        vals = [np.float32(cval) if xm[x + dx] < 0 else src[yy, xm[x + dx]] for dx in range(kx)]
        return _exact_sum(row, vals, contract)

    for (y, x) in ((0, 0), (0, W - 1), (H - 1, 0), (H - 1, W - 1), (1, 8), (5, 2), (6, 9),
                   (H - 2, W - 3), (3, W - 1)):
        hv = [inter(ym[y + dy], x) for dy in range(ky)]
        assert dst[y, x] == _exact_sum(col, hv, contract), (y, x)


@pytest.mark.parametrize("mode", SEP_MODES)
@pytest.mark.parametrize("kx,ky", [(13, 7), (7, 13), (9, 9)])
@pytest.mark.parametrize("contract", [False, True])
def test_separable_float_vs_fp64_correlate1d(mode, kx, ky, contract):
    """The whole image against scipy's FP64 correlate1d along x, then along
    y (the same mode and constant for the intermediate), within the bound of
    float32 sequential summation: |err| <= (g_kx + g_ky + g_kx*g_ky) *
    sum|row| * sum|col| * max|x|, g_n = n u / (1 - n u), u = 2^-24.  A
    swapped or transposed pass misses by orders of magnitude more."""
    H, W = 40, 57
    src = ob.random_image((H, W), np.float32, 200 + kx + ky)
    row = skewed_1d(kx, 31 + kx, signed=True)
    col = skewed_1d(ky, 37 + ky, signed=True)
    cval = -0.625
    dst = np.zeros_like(src)
    ob.separable(src, dst, row, col, boundary=mode, const_val=cval, fp_contract=contract)
    m = SCIPY_MODE[mode]
    ref = ndi.correlate1d(src.astype(np.float64), row.astype(np.float64), axis=1, mode=m,
                          cval=cval)
    ref = ndi.correlate1d(ref, col.astype(np.float64), axis=0, mode=m, cval=cval)
    u = 2.0 ** -24

    def g(n):
        return n * u / (1 - n * u)

    mx = max(float(np.abs(src).max()), abs(cval))
    bound = (g(kx) + g(ky) + g(kx) * g(ky)) * float(np.abs(row).sum()) * \
        float(np.abs(col).sum()) * mx
    err = float(np.abs(dst.astype(np.float64) - ref).max())
    assert err <= bound, (err, bound)
    # the transposed program is far outside the bound
    swapped = np.zeros_like(src)
    if kx != ky:
        return
    ob.separable(src, swapped, col, row, boundary=mode, const_val=cval, fp_contract=contract)
    assert float(np.abs(swapped.astype(np.float64) - ref).max()) > 100 * bound
