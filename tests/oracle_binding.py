"""ctypes binding of the CPU oracle (oracle/liboracle.so) over numpy arrays.

Test infrastructure: builds the same descriptors as hipacc_amd.operators but
with host pointers, and runs the C restatement of the reference operators.
"""
from __future__ import annotations

import ctypes
import subprocess
from ctypes import POINTER, c_double, c_float, c_int, c_uint64, c_void_p
from pathlib import Path
from typing import Optional, Sequence, Tuple

import numpy as np

from hipacc_amd import capi

ROOT = Path(__file__).resolve().parents[1]
ORACLE_DIR = ROOT / "oracle"
ORACLE_LIB = ORACLE_DIR / "liboracle.so"

NP_DTYPES = {capi.U8: np.uint8, capi.I8: np.int8, capi.U16: np.uint16, capi.I16: np.int16,
             capi.U32: np.uint32, capi.I32: np.int32, capi.F32: np.float32}
CODES = {np.dtype(v): k for k, v in NP_DTYPES.items()}

_LIB = None


def lib() -> ctypes.CDLL:
    global _LIB
    if _LIB is None:
        if not ORACLE_LIB.exists():
            subprocess.run(["make", "-C", str(ORACLE_DIR), "liboracle.so"], check=True,
                           capture_output=True)
        L = ctypes.CDLL(str(ORACLE_LIB))
        L.oracle_local.argtypes = [POINTER(capi.LocalDesc)]
        L.oracle_local_f64.argtypes = [POINTER(capi.LocalDesc), c_void_p]
        L.oracle_separable.argtypes = [POINTER(capi.SeparableDesc)]
        L.oracle_point.argtypes = [POINTER(capi.PointDesc)]
        L.oracle_reduce.argtypes = [POINTER(capi.View), c_int, c_int, c_void_p, c_int]
        L.oracle_sum_f64.argtypes = [POINTER(capi.View)]
        L.oracle_sum_f64.restype = c_double
        L.oracle_histogram.argtypes = [POINTER(capi.View), c_int, c_float, c_void_p]
        L.oracle_bilateral.argtypes = [POINTER(capi.BilateralDesc)]
        L.oracle_bilateral_f64.argtypes = [POINTER(capi.BilateralDesc), c_void_p]
        L.oracle_fill_random.argtypes = [POINTER(capi.View), c_uint64]
        L.oracle_num_threads.restype = c_int
        L.oracle_resample.argtypes = [POINTER(capi.ResampleDesc)]
        L.oracle_local_grad.argtypes = [POINTER(capi.LocalGradDesc), c_void_p, c_void_p, c_void_p,
                                        c_void_p]
        L.oracle_bilateral_grad.argtypes = [POINTER(capi.BilateralGradDesc), c_void_p, c_void_p,
                                            c_void_p, c_void_p]
        _LIB = L
    return _LIB


Window = Optional[Tuple[int, int, int, int]]


def hview(a: np.ndarray, window: Window = None, dtype: Optional[int] = None) -> capi.View:
    """View of a (H, W) / (B, H, W) image, or with a vector dtype (capi.U8X4
    ...) of a (H, W, 4) / (B, H, W, 4) image of 4-channel pixels."""
    assert a.flags.c_contiguous or (a.ndim >= 2 and a.strides[-1] == a.itemsize)
    vec = dtype in capi.VECTOR_DTYPES
    if vec:
        assert a.shape[-1] == 4 and a.strides[-1] == a.itemsize
    px = a.itemsize * (4 if vec else 1)      # bytes per pixel
    h, w = (a.shape[-3], a.shape[-2]) if vec else (a.shape[-2], a.shape[-1])
    nd = a.ndim - (1 if vec else 0)
    v = capi.View()
    v.data = a.ctypes.data
    v.dtype = CODES[a.dtype] if dtype is None else dtype
    v.stride = a.strides[-3 if vec else -2] // px
    v.img_width, v.img_height = w, h
    x, y, ww, hh = window if window is not None else (0, 0, w, h)
    v.offset_x, v.offset_y, v.width, v.height = x, y, ww, hh
    v.batch = a.shape[0] if nd == 3 else 1
    v.plane_stride = a.strides[0] // px if nd == 3 else 0
    return v


# ---- synthetic pixels (SplitMix64, identical to the GPU fill kernel) --------
_G = np.uint64(0x9E3779B97F4A7C15)


def splitmix(seed: int, n: int, start: int = 0) -> np.ndarray:
    k = np.arange(start + 1, start + n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + k * _G
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def random_image(shape, dtype, seed: int) -> np.ndarray:
    n = int(np.prod(shape))
    x = splitmix(seed, n)
    dt = np.dtype(dtype)
    if dt == np.uint8:
        v = (x >> np.uint64(56)).astype(np.uint8)
    elif dt == np.int8:
        v = (x >> np.uint64(56)).astype(np.uint8).view(np.int8)
    elif dt == np.uint16:
        v = (x >> np.uint64(48)).astype(np.uint16)
    elif dt == np.int16:
        v = (x >> np.uint64(48)).astype(np.uint16).view(np.int16)
    elif dt == np.uint32:
        v = (x >> np.uint64(32)).astype(np.uint32)
    elif dt == np.int32:
        v = (x >> np.uint64(32)).astype(np.uint32).view(np.int32)
    else:
        v = (x >> np.uint64(40)).astype(np.float32) * np.float32(1.0 / 16777216.0)
    return v.reshape(shape)


def _arr(values, np_dtype):
    a = np.ascontiguousarray(np.asarray(values, dtype=np_dtype).reshape(-1))
    return a, a.ctypes.data


# ---- operators ----------------------------------------------------------------
def local(src: np.ndarray, dst: np.ndarray, mask_w: int, mask_h: int, *, weights=None, domain=None,
          acc_float=True, reduce=capi.SUM, boundary=capi.CLAMP, const_val=0.0,
          cvt=capi.CVT_TRUNC, post_div=0.0, fp_contract=True, src_window: Window = None,
          dst_window: Window = None, dst_dtype=None, f64_out: Optional[np.ndarray] = None,
          src_dtype=None):
    d = capi.LocalDesc()
    d.input = hview(src, src_window, src_dtype)
    d.output = hview(dst, dst_window, dst_dtype)
    d.mask_w, d.mask_h = mask_w, mask_h
    keep = []
    if weights is not None:
        a, p = _arr(weights, np.float32 if acc_float else np.int32)
        if acc_float:
            d.weights_f32 = p
        else:
            d.weights_i32 = p
        keep.append(a)
    d.use_weights = int(weights is not None)
    if domain is not None:
        a, p = _arr(domain, np.uint8)
        d.domain = p
        keep.append(a)
    d.acc_dtype = capi.F32 if acc_float else capi.I32
    d.reduce, d.boundary, d.const_val = reduce, boundary, const_val
    d.cvt, d.post_div, d.fp_contract = cvt, post_div, int(fp_contract)
    if f64_out is not None:
        lib().oracle_local_f64(ctypes.byref(d), f64_out.ctypes.data)
    else:
        lib().oracle_local(ctypes.byref(d))


def separable(src, dst, row, col, *, acc_float=True, boundary=capi.CLAMP, const_val=0.0,
              cvt=capi.CVT_TRUNC, post_div=0.0, fp_contract=True, src_window: Window = None,
              dst_window: Window = None, dst_dtype=None):
    d = capi.SeparableDesc()
    d.input = hview(src, src_window)
    d.output = hview(dst, dst_window, dst_dtype)
    d.ksize_x, d.ksize_y = len(row), len(col)
    ra, rp = _arr(row, np.float32 if acc_float else np.int32)
    ca, cp = _arr(col, np.float32 if acc_float else np.int32)
    if acc_float:
        d.row_f32, d.col_f32 = rp, cp
    else:
        d.row_i32, d.col_i32 = rp, cp
    d.acc_dtype = capi.F32 if acc_float else capi.I32
    d.boundary, d.const_val, d.cvt, d.post_div = boundary, const_val, cvt, post_div
    d.fp_contract = int(fp_contract)
    lib().oracle_separable(ctypes.byref(d))


def point(op, src0, dst, src1=None, *, alpha=1.0, beta=0.0, cvt=capi.CVT_TRUNC,
          windows=(None, None, None), dst_dtype=None):
    d = capi.PointDesc()
    d.in0 = hview(src0, windows[0])
    if src1 is not None:
        d.in1 = hview(src1, windows[1])
    d.output = hview(dst, windows[2], dst_dtype)
    d.op, d.alpha, d.beta, d.cvt = op, alpha, beta, cvt
    lib().oracle_point(ctypes.byref(d))


def chain(src: np.ndarray, stages, *, src_window: Window = None) -> np.ndarray:
    """The unfused chain of point and local stages (the stage dicts of
    hipacc_amd.operators.prepare_fused, with numpy ``in1``): each stage an
    oracle call over the previous stage's image; returns the last image."""
    cur, win = src, src_window
    shape = src.shape if src_window is None else src.shape[:-2] + (src_window[3], src_window[2])
    for k, st in enumerate(stages):
        st = dict(st)
        kind = st.pop("kind")
        nxt = np.zeros(shape, NP_DTYPES[st.pop("dtype")])
        if kind == "point":
            in1 = st.pop("in1", None)
            point(st.pop("op", capi.POINT_AFFINE), cur, nxt, in1, alpha=st.pop("alpha", 1.0),
                  beta=st.pop("beta", 0.0), cvt=st.pop("cvt", capi.CVT_TRUNC),
                  windows=(win, st.pop("in1_window", None), None))
        else:
            local(cur, nxt, st.pop("mask_w"), st.pop("mask_h"), src_window=win, **st)
        cur, win = nxt, None
    return cur


def reduce(src, mode, acc_dtype=None, window: Window = None, nthreads: int = 1):
    v = hview(src, window)
    if acc_dtype is None:
        acc_dtype = capi.F32 if v.dtype == capi.F32 else capi.U32
    out = np.zeros(1, dtype=NP_DTYPES[acc_dtype])
    lib().oracle_reduce(ctypes.byref(v), mode, acc_dtype, out.ctypes.data, nthreads)
    return out[0]


def sum_f64(src, window: Window = None) -> float:
    v = hview(src, window)
    return lib().oracle_sum_f64(ctypes.byref(v))


def histogram(src, num_bins, scale=1.0, window: Window = None):
    v = hview(src, window)
    out = np.zeros(num_bins, dtype=np.uint32)
    lib().oracle_histogram(ctypes.byref(v), num_bins, scale, out.ctypes.data)
    return out


def bilateral(src, dst, ksize, sigma_d, sigma_r, *, spatial=None, boundary=capi.MIRROR,
              const_val=0.0, cvt=capi.CVT_TRUNC, f64_out: Optional[np.ndarray] = None):
    d = capi.BilateralDesc()
    d.input = hview(src)
    d.output = hview(dst)
    d.ksize, d.sigma_d, d.sigma_r = ksize, sigma_d, sigma_r
    keep = None
    if spatial is not None:
        keep, d.spatial = _arr(spatial, np.float32)
    d.boundary, d.const_val, d.cvt = boundary, const_val, cvt
    if f64_out is not None:
        lib().oracle_bilateral_f64(ctypes.byref(d), f64_out.ctypes.data)
    else:
        lib().oracle_bilateral(ctypes.byref(d))
    del keep


# ---- adjoints ----------------------------------------------------------------
def local_grad(src, grad_out, weights, mask_w, mask_h, *, boundary=capi.CLAMP, const_val=0.0,
               with_abs=False):
    """(dx, dw[, dx_abs, dw_abs]) of the float weighted SUM, in double."""
    d = capi.LocalGradDesc()
    d.input = hview(src)
    d.grad_out = hview(grad_out)
    d.mask_w, d.mask_h = mask_w, mask_h
    w, d.weights_f32 = _arr(weights, np.float32)
    d.boundary, d.const_val = boundary, const_val
    dx = np.zeros(src.shape, np.float64)
    dw = np.zeros(mask_w * mask_h, np.float64)
    dxa = np.zeros_like(dx) if with_abs else None
    dwa = np.zeros_like(dw) if with_abs else None
    lib().oracle_local_grad(ctypes.byref(d), dx.ctypes.data, dw.ctypes.data,
                            dxa.ctypes.data if with_abs else None,
                            dwa.ctypes.data if with_abs else None)
    del w
    return (dx, dw, dxa, dwa) if with_abs else (dx, dw)


def bilateral_grad(src, grad_out, ksize, sigma_d, sigma_r, *, spatial=None, boundary=capi.MIRROR,
                   const_val=0.0, with_abs=False):
    """(dx, dsigma[, dx_abs, dsigma_abs]) of the bilateral filter, in double."""
    d = capi.BilateralGradDesc()
    d.input = hview(src)
    d.grad_out = hview(grad_out)
    d.ksize, d.sigma_d, d.sigma_r = ksize, sigma_d, sigma_r
    keep = None
    if spatial is not None:
        keep, d.spatial = _arr(spatial, np.float32)
    d.boundary, d.const_val = boundary, const_val
    dx = np.zeros(src.shape, np.float64)
    ds = np.zeros(1, np.float64)
    dxa = np.zeros_like(dx) if with_abs else None
    dsa = np.zeros(1, np.float64) if with_abs else None
    lib().oracle_bilateral_grad(ctypes.byref(d), dx.ctypes.data, ds.ctypes.data,
                                dxa.ctypes.data if with_abs else None,
                                dsa.ctypes.data if with_abs else None)
    del keep
    return (dx, ds[0], dxa, dsa[0]) if with_abs else (dx, ds[0])


# ---- resampling and pyramids ---------------------------------------------------
def resample(src, dst, interp, *, op=capi.RESAMPLE_COPY, src2=None, boundary=capi.CLAMP,
             const_val=0.0, src_window: Window = None, dst_window: Window = None, xf=0, yf=0):
    d = capi.ResampleDesc()
    d.input = hview(src, src_window)
    if src2 is not None:
        d.in2 = hview(src2)
    d.output = hview(dst, dst_window)
    d.interp, d.op, d.boundary, d.const_val = interp, op, boundary, const_val
    d.xf, d.yf = xf, yf
    if lib().oracle_resample(ctypes.byref(d)):
        raise ValueError(f"oracle_resample: interp {interp}")


BINOMIAL5 = (np.outer([1, 4, 6, 4, 1], [1, 4, 6, 4, 1]) / 256.0).astype(np.float32).ravel()


def pyramid(g0, levels, *, mask=BINOMIAL5, boundary=capi.CLAMP, fp_contract=True):
    """Gaussian and Laplacian pyramids of float image g0 by the pyramid
    sample's per-level kernel sequence: T = conv5x5(G[l]); G[l+1] = NN(T);
    L[l] = G[l] - LF(G[l+1]).  Returns (G list, L list of levels-1)."""
    G, L = [np.ascontiguousarray(g0, np.float32)], []
    for _ in range(levels - 1):
        g = G[-1]
        h, w = g.shape
        t = np.zeros_like(g)
        local(g, t, 5, 5, weights=mask, boundary=boundary, fp_contract=fp_contract)
        g1 = np.zeros((h // 2, w // 2), np.float32)
        resample(t, g1, capi.INTERP_NN)
        lap = np.zeros_like(g)
        resample(g1, lap, capi.INTERP_LF, op=capi.RESAMPLE_SUB, src2=g)
        G.append(g1)
        L.append(lap)
    return G, L
