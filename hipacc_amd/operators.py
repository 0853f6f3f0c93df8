"""Host-side calls of the C ABI on torch device tensors.

Images are 2D tensors (H, W) -- row pitch taken from ``tensor.stride(0)`` -- or
3D tensors (B, H, W) for batches.  Each function builds the descriptor of
include/hipacc_hip.h and launches on torch's current stream.  Windows
(accessor regions / iteration spaces) are given as ``(x, y, w, h)``; the
default is the whole image.

The operator vocabulary is Hipacc's: ``local`` is Kernel::convolve /
reduce(Domain) / iterate, ``point`` is a point operator's output(),
``reduce`` is reduced_data(), ``histogram`` is binned_data() of a counting
binning, ``bilateral`` is the Bilateral_Filter operator.
"""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence, Tuple

import numpy as np
import torch

from . import capi

_TORCH_DTYPES = {
    torch.uint8: capi.U8, torch.int8: capi.I8, torch.int16: capi.I16, torch.int32: capi.I32,
    torch.float32: capi.F32,
}
for _name, _code in (("uint16", capi.U16), ("uint32", capi.U32)):
    if hasattr(torch, _name):
        _TORCH_DTYPES[getattr(torch, _name)] = _code

Window = Optional[Tuple[int, int, int, int]]


def dtype_code(t: torch.Tensor) -> int:
    try:
        return _TORCH_DTYPES[t.dtype]
    except KeyError:
        raise TypeError(f"unsupported pixel dtype {t.dtype}") from None


def view(t: torch.Tensor, window: Window = None, dtype: Optional[int] = None) -> capi.View:
    """hipacc_view of a (H, W) or (B, H, W) device tensor, rows contiguous; with
    a vector dtype (capi.U8X4, I16X4, I32X4, F32X4) of a (H, W, 4) or
    (B, H, W, 4) tensor of 4-channel pixels."""
    if not t.is_cuda:
        raise ValueError("image tensors must live on the GPU")
    vec = dtype in capi.VECTOR_DTYPES
    nd = t.dim() - (1 if vec else 0)
    if nd not in (2, 3) or t.stride(-1) != 1 or (vec and (t.shape[-1] != 4 or t.stride(-2) != 4)):
        raise ValueError("images are (H, W) or (B, H, W) tensors with unit column stride "
                         "((..., 4) with interleaved channels for vector pixels)")
    lanes = 4 if vec else 1
    h, w = (t.shape[-3], t.shape[-2]) if vec else (t.shape[-2], t.shape[-1])
    v = capi.View()
    v.data = t.data_ptr()
    v.dtype = dtype_code(t) if dtype is None else dtype
    v.stride = t.stride(-3 if vec else -2) // lanes
    v.img_width, v.img_height = w, h
    x, y, ww, hh = window if window is not None else (0, 0, w, h)
    v.offset_x, v.offset_y, v.width, v.height = x, y, ww, hh
    v.batch = t.shape[0] if nd == 3 else 1
    v.plane_stride = t.stride(0) // lanes if nd == 3 else 0
    return v


def _stream() -> int:
    return torch.cuda.current_stream().cuda_stream


class Call:
    """A prepared C-ABI call: the descriptor is built once (host tables kept
    alive here) and ``launch(stream)`` only crosses the ABI.  ``variant``
    forces a kernel variant (capi.VAR_*) with its tile parameter."""

    _VARIANT_FN = {"hipacc_hip_local": "hipacc_hip_local_variant",
                   "hipacc_hip_separable": "hipacc_hip_separable_variant",
                   "hipacc_hip_local_grad": "hipacc_hip_local_grad_variant"}
    _TUNE_FN = {"hipacc_hip_local": "hipacc_hip_tune_local",
                "hipacc_hip_separable": "hipacc_hip_tune_separable"}

    def __init__(self, fn_name: str, desc, keep=(), variant: int = 0, param: int = 0):
        self.name = fn_name
        self.desc = desc
        self.ref = ctypes.byref(desc)
        self.keep = list(keep)
        self.variant, self.param = variant, param
        if variant:
            self.fn = getattr(capi.lib(), self._VARIANT_FN[fn_name])
            self.extra = (variant,) if fn_name == "hipacc_hip_local_grad" else (variant, param)
        else:
            self.fn = getattr(capi.lib(), fn_name)
            self.extra = ()

    def launch(self, stream: Optional[int] = None) -> None:
        st = self.fn(self.ref, *self.extra, _stream() if stream is None else stream)
        if st:
            capi.check(st, self.name)

    def tune(self, iters: int = 20) -> capi.TuneEntry:
        """Times every kernel variant of this operator on its buffers and
        records the fastest in the library's tune cache."""
        e = capi.TuneEntry()
        fn = getattr(capi.lib(), self._TUNE_FN[self.name])
        capi.check(fn(self.ref, iters, ctypes.byref(e), _stream()), self._TUNE_FN[self.name])
        return e

    def with_variant(self, variant: int, param: int = 0) -> "Call":
        return Call(self.name, self.desc, self.keep, variant, param)


def _arr(values, np_dtype):
    a = np.ascontiguousarray(np.asarray(values, dtype=np_dtype).reshape(-1))
    return a, a.ctypes.data


def local(*args, **kw) -> None:
    prepare_local(*args, **kw).launch()


def prepare_local(src: torch.Tensor, dst: torch.Tensor, mask_w: int, mask_h: int, *,
          weights: Optional[Sequence[float]] = None, weights_dev: Optional[torch.Tensor] = None,
          domain: Optional[Sequence[int]] = None, acc_float: bool = True,
          reduce: int = capi.SUM, boundary: int = capi.CLAMP, const_val: float = 0.0,
          cvt: int = capi.CVT_TRUNC, post_div: float = 0.0, fp_contract: bool = True,
          src_window: Window = None, dst_window: Window = None,
          dst_dtype: Optional[int] = None, src_dtype: Optional[int] = None) -> Call:
    d = capi.LocalDesc()
    d.input = view(src, src_window, src_dtype)
    d.output = view(dst, dst_window, dst_dtype)
    d.mask_w, d.mask_h = mask_w, mask_h
    keep = []
    if weights is not None:
        if acc_float:
            a, p = _arr(weights, np.float32)
            d.weights_f32 = p
        else:
            a, p = _arr(weights, np.int32)
            d.weights_i32 = p
        keep.append(a)
    if weights_dev is not None:
        d.weights_dev = weights_dev.data_ptr()
    d.use_weights = int(weights is not None or weights_dev is not None)
    if domain is not None:
        a, p = _arr(domain, np.uint8)
        d.domain = p
        keep.append(a)
    d.acc_dtype = capi.F32 if acc_float else capi.I32
    d.reduce, d.boundary, d.const_val = reduce, boundary, const_val
    d.cvt, d.post_div, d.fp_contract = cvt, post_div, int(fp_contract)
    return Call("hipacc_hip_local", d, keep)


def separable(*args, **kw) -> None:
    prepare_separable(*args, **kw).launch()


def prepare_separable(src: torch.Tensor, dst: torch.Tensor, row: Sequence, col: Sequence, *,
              acc_float: bool = True, boundary: int = capi.CLAMP, const_val: float = 0.0,
              cvt: int = capi.CVT_TRUNC, post_div: float = 0.0, fp_contract: bool = True,
              src_window: Window = None, dst_window: Window = None,
              dst_dtype: Optional[int] = None, src_dtype: Optional[int] = None) -> Call:
    d = capi.SeparableDesc()
    d.input = view(src, src_window, src_dtype)
    d.output = view(dst, dst_window, dst_dtype)
    d.ksize_x, d.ksize_y = len(row), len(col)
    if acc_float:
        ra, rp = _arr(row, np.float32)
        ca, cp = _arr(col, np.float32)
        d.row_f32, d.col_f32 = rp, cp
    else:
        ra, rp = _arr(row, np.int32)
        ca, cp = _arr(col, np.int32)
        d.row_i32, d.col_i32 = rp, cp
    d.acc_dtype = capi.F32 if acc_float else capi.I32
    d.boundary, d.const_val, d.cvt, d.post_div = boundary, const_val, cvt, post_div
    d.fp_contract = int(fp_contract)
    return Call("hipacc_hip_separable", d, (ra, ca))


def point(*args, **kw) -> None:
    prepare_point(*args, **kw).launch()


def prepare_point(op: int, src0: torch.Tensor, dst: torch.Tensor,
                  src1: Optional[torch.Tensor] = None, *,
          alpha: float = 1.0, beta: float = 0.0, cvt: int = capi.CVT_TRUNC,
          windows: Tuple[Window, Window, Window] = (None, None, None),
          dst_dtype: Optional[int] = None, src_dtype: Optional[int] = None) -> Call:
    d = capi.PointDesc()
    d.in0 = view(src0, windows[0], src_dtype)
    if src1 is not None:
        d.in1 = view(src1, windows[1], src_dtype)
    d.output = view(dst, windows[2], dst_dtype)
    d.op, d.alpha, d.beta, d.cvt = op, alpha, beta, cvt
    return Call("hipacc_hip_point", d)


def fused(*args, **kw) -> None:
    prepare_fused(*args, **kw).launch()


def prepare_fused(src: torch.Tensor, dst: torch.Tensor, stages, *, src_window: Window = None,
                  dst_window: Window = None, src_dtype: Optional[int] = None,
                  dst_dtype: Optional[int] = None) -> Call:
    """hipacc_hip_fused: a chain of point and local stages in one launch.

    Each stage is a dict with ``kind`` ("point" or "local") and ``dtype``
    (capi pixel type of the image it writes).  A point stage takes ``op``
    (default POINT_AFFINE), ``alpha``, ``beta``, ``cvt`` and, for two-input
    ops, ``in1`` (device tensor) with ``in1_window``; a local stage takes
    prepare_local's keyword arguments (mask_w, mask_h, weights, domain,
    acc_float, reduce, boundary, const_val, cvt, post_div, fp_contract)."""
    if not 1 <= len(stages) <= capi.MAX_FUSED_STAGES:
        raise ValueError(f"1..{capi.MAX_FUSED_STAGES} stages")
    d = capi.FusedDesc()
    d.input = view(src, src_window, src_dtype)
    d.output = view(dst, dst_window, dst_dtype)
    d.num_stages = len(stages)
    keep = []
    for k, st in enumerate(stages):
        st = dict(st)
        g = d.stages[k]
        kind = st.pop("kind")
        g.dtype = st.pop("dtype")
        if kind == "point":
            g.kind = capi.STAGE_POINT
            g.op = st.pop("op", capi.POINT_AFFINE)
            g.alpha, g.beta = st.pop("alpha", 1.0), st.pop("beta", 0.0)
            g.cvt = st.pop("cvt", capi.CVT_TRUNC)
            in1 = st.pop("in1", None)
            win1 = st.pop("in1_window", None)
            if in1 is not None:
                g.in1 = view(in1, win1, st.pop("in1_dtype", None))
            if st:
                raise ValueError(f"point stage {k}: unknown keys {sorted(st)}")
        elif kind == "local":
            g.kind = capi.STAGE_LOCAL
            lc = prepare_local(src, dst, st.pop("mask_w"), st.pop("mask_h"), **st)
            keep.append(lc)
            g.local = ctypes.pointer(lc.desc)
        else:
            raise ValueError(f"stage {k}: kind {kind!r}")
    return Call("hipacc_hip_fused", d, keep)


class ReduceCall:
    """Prepared hipacc_hip_reduce (reduced_data() into a device scalar)."""

    def __init__(self, src, mode, result, acc_dtype=None, window=None, workspace=None):
        self.v = view(src, window)
        if acc_dtype is None:
            acc_dtype = capi.F32 if self.v.dtype == capi.F32 else capi.U32
        self.args = (ctypes.byref(self.v), mode, acc_dtype, result.data_ptr(),
                     workspace.data_ptr() if workspace is not None else None)
        self.fn = capi.lib().hipacc_hip_reduce

    def launch(self, stream: Optional[int] = None) -> None:
        st = self.fn(*self.args, _stream() if stream is None else stream)
        if st:
            capi.check(st, "hipacc_hip_reduce")


def reduce(src: torch.Tensor, mode: int, result: torch.Tensor, *, acc_dtype: Optional[int] = None,
           window: Window = None, workspace: Optional[torch.Tensor] = None) -> None:
    """Writes reduced_data() of the window into the one-element device tensor
    `result` (asynchronous)."""
    ReduceCall(src, mode, result, acc_dtype, window, workspace).launch()


class HistogramCall:
    """Prepared hipacc_hip_histogram."""

    def __init__(self, src, num_bins, bins, scale=1.0, window=None):
        self.v = view(src, window)
        self.args = (ctypes.byref(self.v), num_bins, scale, bins.data_ptr())
        self.fn = capi.lib().hipacc_hip_histogram

    def launch(self, stream: Optional[int] = None) -> None:
        st = self.fn(*self.args, _stream() if stream is None else stream)
        if st:
            capi.check(st, "hipacc_hip_histogram")


def histogram(src: torch.Tensor, num_bins: int, bins: torch.Tensor, *, scale: float = 1.0,
              window: Window = None) -> None:
    HistogramCall(src, num_bins, bins, scale, window).launch()


def bilateral(*args, **kw) -> None:
    prepare_bilateral(*args, **kw).launch()


def prepare_bilateral(src: torch.Tensor, dst: torch.Tensor, ksize: int, sigma_d: float,
                      sigma_r: float, *,
              spatial: Optional[Sequence[float]] = None, boundary: int = capi.MIRROR,
              const_val: float = 0.0, cvt: int = capi.CVT_TRUNC, src_window: Window = None,
              dst_window: Window = None) -> Call:
    d = capi.BilateralDesc()
    d.input = view(src, src_window)
    d.output = view(dst, dst_window)
    d.ksize, d.sigma_d, d.sigma_r = ksize, sigma_d, sigma_r
    keep = None
    if spatial is not None:
        keep, d.spatial = _arr(spatial, np.float32)
    d.boundary, d.const_val, d.cvt = boundary, const_val, cvt
    return Call("hipacc_hip_bilateral", d, [keep])


def prepare_local_grad(src: torch.Tensor, grad_out: torch.Tensor,
                       grad_in: Optional[torch.Tensor], mask_w: int, mask_h: int, *,
                       weights: Optional[Sequence[float]] = None,
                       weights_dev: Optional[torch.Tensor] = None,
                       grad_weights: Optional[torch.Tensor] = None, boundary: int = capi.CLAMP,
                       const_val: float = 0.0,
                       workspace: Optional[torch.Tensor] = None) -> Call:
    """Adjoint of the float weighted-SUM local operator: dL/dX into grad_in
    and dL/dw into grad_weights (a device float tensor of mask_h*mask_w)."""
    d = capi.LocalGradDesc()
    d.input = view(src)
    d.grad_out = view(grad_out)
    if grad_in is not None:
        d.grad_in = view(grad_in)
    d.mask_w, d.mask_h = mask_w, mask_h
    keep = []
    if weights_dev is not None:
        d.weights_dev = weights_dev.data_ptr()
        keep.append(weights_dev)
    elif weights is not None:
        a, d.weights_f32 = _arr(weights, np.float32)
        keep.append(a)
    if grad_weights is not None:
        d.grad_weights_dev = grad_weights.data_ptr()
    d.boundary, d.const_val = boundary, const_val
    if workspace is not None:
        d.workspace = workspace.data_ptr()
    return Call("hipacc_hip_local_grad", d, keep)


def _bilateral_grad_desc(src, dst, norm, ksize, sigma_d, sigma_r, spatial, boundary, const_val):
    d = capi.BilateralGradDesc()
    d.input = view(src)
    d.output = view(dst)
    d.norm = view(norm)
    d.ksize, d.sigma_d, d.sigma_r = ksize, sigma_d, sigma_r
    keep = []
    if spatial is not None:
        a, d.spatial = _arr(spatial, np.float32)
        keep.append(a)
    d.boundary, d.const_val = boundary, const_val
    return d, keep


def prepare_bilateral_train(src: torch.Tensor, dst: torch.Tensor, norm: torch.Tensor, ksize: int,
                            sigma_d: float, sigma_r: float, *,
                            spatial: Optional[Sequence[float]] = None,
                            boundary: int = capi.MIRROR, const_val: float = 0.0) -> Call:
    """Bilateral forward that also writes the per-pixel normaliser."""
    d, keep = _bilateral_grad_desc(src, dst, norm, ksize, sigma_d, sigma_r, spatial, boundary,
                                   const_val)
    return Call("hipacc_hip_bilateral_train", d, keep)


def prepare_bilateral_grad(src: torch.Tensor, dst: torch.Tensor, norm: torch.Tensor,
                           grad_out: torch.Tensor, grad_in: Optional[torch.Tensor], ksize: int,
                           sigma_d: float, sigma_r: float, *,
                           grad_sigma: Optional[torch.Tensor] = None,
                           spatial: Optional[Sequence[float]] = None,
                           boundary: int = capi.MIRROR, const_val: float = 0.0,
                           workspace: Optional[torch.Tensor] = None) -> Call:
    """Bilateral backward: dL/dX into grad_in, dL/dsigma_r into grad_sigma
    (a one-element device float tensor)."""
    d, keep = _bilateral_grad_desc(src, dst, norm, ksize, sigma_d, sigma_r, spatial, boundary,
                                   const_val)
    d.grad_out = view(grad_out)
    if grad_in is not None:
        d.grad_in = view(grad_in)
    if grad_sigma is not None:
        d.grad_sigma_r_dev = grad_sigma.data_ptr()
    if workspace is not None:
        d.workspace = workspace.data_ptr()
    return Call("hipacc_hip_bilateral_grad", d, keep)


def prepare_resample(src: torch.Tensor, dst: torch.Tensor, interp: int, *,
                     op: int = capi.RESAMPLE_COPY, src2: Optional[torch.Tensor] = None,
                     boundary: int = capi.CLAMP, const_val: float = 0.0,
                     src_window: Window = None, dst_window: Window = None, xf: int = 0,
                     yf: int = 0) -> Call:
    """Point operator reading `src` (f32 or u8; dst of the same type)
    through Interpolate::NN / LF / B5 / CF / L3 at the iteration-space pixel
    offset (xf, yf); op SUB gives src2 - v, ADD v + src2 (f32)."""
    d = capi.ResampleDesc()
    d.input = view(src, src_window)
    if src2 is not None:
        d.in2 = view(src2)
    d.output = view(dst, dst_window)
    d.interp, d.op, d.boundary, d.const_val = interp, op, boundary, const_val
    d.xf, d.yf = xf, yf
    return Call("hipacc_hip_resample", d, [src2])


def resample(*args, **kw) -> None:
    prepare_resample(*args, **kw).launch()


def pyramid_shapes(h: int, w: int, levels: int):
    """Level sizes of a pyramid: halved with truncation."""
    return [(h >> l, w >> l) for l in range(levels)]


def pyramid_plan_tiles(shapes, levels: int, tile_rows: int = 0):
    """Narrow and wide tile counts per level of the work lists a pyramid plan
    builds for level-0 shapes [(h, w), ...]; host only, no device call.
    Raises when a tile would not fit its kernel's LDS capacity."""
    n = len(shapes)
    hs = np.array([h for h, _ in shapes], np.int32)
    ws = np.array([w for _, w in shapes], np.int32)
    narrow = np.zeros(max(levels - 1, 1), np.int32)
    wide = np.zeros(max(levels - 1, 1), np.int32)
    capi.check(capi.lib().hipacc_hip_pyramid_plan_tiles(
        ws.ctypes.data, hs.ctypes.data, n, levels, tile_rows, narrow.ctypes.data,
        wide.ctypes.data), "hipacc_hip_pyramid_plan_tiles")
    return narrow[:levels - 1].tolist(), wide[:levels - 1].tolist()


class PyramidPlan:
    """Batched Gaussian / Laplacian pyramids of many images.

    gauss[i][l] / lap[i][l] are float32 (H_l, W_l) device tensors (gauss[i][0]
    is image i's input; lap[i] has levels-1 entries).  ``run()`` replays the
    captured hipGraph, ``run_direct()`` issues the same launches without it."""

    def __init__(self, gauss, lap, *, mask=None, boundary: int = capi.CLAMP,
                 fp_contract: bool = True, tile_rows: int = 0):
        levels = len(gauss[0])
        n = len(gauss)
        self.images = (capi.PyramidImage * n)()
        for i in range(n):
            im = self.images[i]
            im.height, im.width = gauss[i][0].shape
            for l in range(levels):
                g = gauss[i][l]
                if g.dtype != torch.float32 or not g.is_cuda or g.stride(-1) != 1:
                    raise ValueError("pyramid levels must be float32 device tensors with unit "
                                     "column stride")
                im.gauss[l] = g.data_ptr()
                im.gauss_stride[l] = g.stride(0)
                if l < levels - 1:
                    im.lap[l] = lap[i][l].data_ptr()
                    im.lap_stride[l] = lap[i][l].stride(0)
        d = capi.PyramidDesc()
        d.num_images, d.levels = n, levels
        d.images = self.images
        self._mask = None
        if mask is not None:
            self._mask, d.mask = _arr(mask, np.float32)
        d.boundary, d.fp_contract, d.tile_rows = boundary, int(fp_contract), tile_rows
        self._keep = (gauss, lap)
        self.plan = ctypes.c_void_p()
        capi.check(capi.lib().hipacc_hip_pyramid_plan_create(ctypes.byref(d), _stream(),
                                                              ctypes.byref(self.plan)),
                   "hipacc_hip_pyramid_plan_create")

    def run(self, stream: Optional[int] = None) -> None:
        capi.check(capi.lib().hipacc_hip_pyramid_run(self.plan, _stream() if stream is None
                                                     else stream), "hipacc_hip_pyramid_run")

    def run_direct(self, stream: Optional[int] = None) -> None:
        capi.check(capi.lib().hipacc_hip_pyramid_run_direct(
            self.plan, _stream() if stream is None else stream), "hipacc_hip_pyramid_run_direct")

    def close(self) -> None:
        if self.plan:
            capi.lib().hipacc_hip_pyramid_plan_destroy(self.plan)
            self.plan = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def fill_random(t: torch.Tensor, seed: int, window: Window = None) -> None:
    v = view(t, window)
    capi.check(capi.lib().hipacc_hip_fill_random(ctypes.byref(v), seed, _stream()),
               "hipacc_hip_fill_random")


def init(device: int = 0) -> None:
    capi.check(capi.lib().hipacc_hip_init(device), "hipacc_hip_init")
