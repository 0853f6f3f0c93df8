"""ctypes binding of libhipacc_gfx950.so (include/hipacc_hip.h).

The structures below mirror the C declarations field for field.  ``lib()``
loads the in-tree library and raises if it is missing: there is no CPU
fallback anywhere in this package.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_double, c_float, c_int32, c_int64, c_size_t, c_uint32, c_uint64, c_void_p
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libhipacc_gfx950.so"

# ---- enums (values as in include/hipacc_hip.h) -----------------------------
U8, I8, U16, I16, U32, I32, F32 = range(7)
# 4-wide vector pixels (uchar4, short4, int4, float4): arrays / tensors of
# shape (..., H, W, 4), channels interleaved
U8X4, I16X4, I32X4, F32X4 = 7, 8, 9, 10
VECTOR_DTYPES = (U8X4, I16X4, I32X4, F32X4)
DTYPE_BYTES = {U8: 1, I8: 1, U16: 2, I16: 2, U32: 4, I32: 4, F32: 4}
UNDEFINED, CLAMP, REPEAT, MIRROR, CONSTANT = range(5)
BOUNDARY_NAMES = {UNDEFINED: "UNDEFINED", CLAMP: "CLAMP", REPEAT: "REPEAT", MIRROR: "MIRROR",
                  CONSTANT: "CONSTANT"}
SUM, MIN, MAX, PROD = range(4)
CVT_TRUNC, CVT_ROUND = 0, 1
(POINT_COPY, POINT_AFFINE, POINT_ADD, POINT_SUB, POINT_MUL, POINT_ABSDIFF, POINT_MIN, POINT_MAX,
 POINT_MAGNITUDE) = range(9)
MAX_MASK_DIM = 15
MAX_SEP_RADIUS = 12

OK = 0


class View(ctypes.Structure):
    _fields_ = [("data", c_void_p), ("dtype", c_int32), ("stride", c_int32),
                ("img_width", c_int32), ("img_height", c_int32), ("width", c_int32),
                ("height", c_int32), ("offset_x", c_int32), ("offset_y", c_int32),
                ("batch", c_int32), ("plane_stride", c_int64)]


class LocalDesc(ctypes.Structure):
    _fields_ = [("input", View), ("output", View), ("mask_w", c_int32), ("mask_h", c_int32),
                ("weights_f32", c_void_p), ("weights_i32", c_void_p), ("weights_dev", c_void_p),
                ("domain", c_void_p), ("use_weights", c_int32), ("acc_dtype", c_int32),
                ("reduce", c_int32), ("boundary", c_int32), ("const_val", c_double),
                ("cvt", c_int32), ("post_div", c_float), ("fp_contract", c_int32)]


MAX_FUSED_STAGES = 8
MAX_FUSED_LOCALS = 2
STAGE_POINT, STAGE_LOCAL = 0, 1


class FusedStage(ctypes.Structure):
    _fields_ = [("kind", c_int32), ("dtype", c_int32), ("op", c_int32), ("alpha", c_float),
                ("beta", c_float), ("cvt", c_int32), ("in1", View),
                ("local", POINTER(LocalDesc))]


class FusedDesc(ctypes.Structure):
    _fields_ = [("input", View), ("output", View), ("num_stages", c_int32),
                ("stages", FusedStage * MAX_FUSED_STAGES)]


class SeparableDesc(ctypes.Structure):
    _fields_ = [("input", View), ("output", View), ("ksize_x", c_int32), ("ksize_y", c_int32),
                ("row_f32", c_void_p), ("col_f32", c_void_p), ("row_i32", c_void_p),
                ("col_i32", c_void_p), ("acc_dtype", c_int32), ("boundary", c_int32),
                ("const_val", c_double), ("cvt", c_int32), ("post_div", c_float),
                ("fp_contract", c_int32)]


class PointDesc(ctypes.Structure):
    _fields_ = [("in0", View), ("in1", View), ("output", View), ("op", c_int32),
                ("alpha", c_float), ("beta", c_float), ("cvt", c_int32)]


class BilateralDesc(ctypes.Structure):
    _fields_ = [("input", View), ("output", View), ("ksize", c_int32), ("sigma_d", c_float),
                ("sigma_r", c_float), ("spatial", c_void_p), ("boundary", c_int32),
                ("const_val", c_double), ("cvt", c_int32)]


class LocalGradDesc(ctypes.Structure):
    _fields_ = [("input", View), ("grad_out", View), ("grad_in", View), ("mask_w", c_int32),
                ("mask_h", c_int32), ("weights_f32", c_void_p), ("weights_dev", c_void_p),
                ("grad_weights_dev", c_void_p), ("boundary", c_int32), ("const_val", c_double),
                ("workspace", c_void_p)]


class BilateralGradDesc(ctypes.Structure):
    _fields_ = [("input", View), ("output", View), ("norm", View), ("grad_out", View),
                ("grad_in", View), ("ksize", c_int32), ("sigma_d", c_float), ("sigma_r", c_float),
                ("spatial", c_void_p), ("boundary", c_int32), ("const_val", c_double),
                ("grad_sigma_r_dev", c_void_p), ("workspace", c_void_p)]


INTERP_NN, INTERP_LF, INTERP_B5, INTERP_CF, INTERP_L3 = 1, 2, 3, 4, 5
RESAMPLE_COPY, RESAMPLE_SUB, RESAMPLE_ADD = 0, 1, 2
MAX_LEVELS = 16


class ResampleDesc(ctypes.Structure):
    _fields_ = [("input", View), ("in2", View), ("output", View), ("interp", c_int32),
                ("op", c_int32), ("boundary", c_int32), ("const_val", c_double),
                ("xf", c_int32), ("yf", c_int32)]


class PyramidImage(ctypes.Structure):
    _fields_ = [("width", c_int32), ("height", c_int32),
                ("gauss", c_void_p * MAX_LEVELS), ("gauss_stride", c_int32 * MAX_LEVELS),
                ("lap", c_void_p * MAX_LEVELS), ("lap_stride", c_int32 * MAX_LEVELS)]


class PyramidDesc(ctypes.Structure):
    _fields_ = [("num_images", c_int32), ("levels", c_int32),
                ("images", POINTER(PyramidImage)), ("mask", c_void_p), ("boundary", c_int32),
                ("fp_contract", c_int32), ("tile_rows", c_int32)]


class TuneEntry(ctypes.Structure):
    _fields_ = [("kind", c_int32), ("kx", c_int32), ("ky", c_int32), ("in_dtype", c_int32),
                ("out_dtype", c_int32), ("acc_dtype", c_int32), ("boundary", c_int32),
                ("fp_contract", c_int32), ("weights_dev", c_int32), ("size_class", c_int32),
                ("variant", c_int32),
                ("param", c_int32), ("ms", c_float)]


NOALLOC_SCRATCH_BYTES = 16 * 1024 * 1024   # HIPACC_NOALLOC_SCRATCH_BYTES

VAR_AUTO, VAR_TILE, VAR_STREAM, VAR_GENERIC, VAR_WIDE, VAR_RING = range(6)
VARIANT_NAMES = {VAR_TILE: "tile", VAR_STREAM: "stream", VAR_GENERIC: "generic", VAR_WIDE: "wide",
                 VAR_RING: "ring"}

# Every exported entry point with its ctypes signature; tests check that the
# library exports exactly these names.
SIGNATURES = {
    "hipacc_hip_init": (c_int32, [c_int32]),
    "hipacc_hip_strerror": (ctypes.c_char_p, [c_int32]),
    "hipacc_hip_last_error": (ctypes.c_char_p, []),
    "hipacc_hip_version": (ctypes.c_char_p, []),
    "hipacc_hip_local": (c_int32, [POINTER(LocalDesc), c_void_p]),
    "hipacc_hip_separable": (c_int32, [POINTER(SeparableDesc), c_void_p]),
    "hipacc_hip_point": (c_int32, [POINTER(PointDesc), c_void_p]),
    "hipacc_hip_fused": (c_int32, [POINTER(FusedDesc), c_void_p]),
    "hipacc_hip_reduce_workspace_bytes": (c_size_t, []),
    "hipacc_hip_reduce": (c_int32, [POINTER(View), c_int32, c_int32, c_void_p, c_void_p, c_void_p]),
    "hipacc_hip_reduce_sync": (c_int32, [POINTER(View), c_int32, c_int32, c_void_p, c_void_p]),
    "hipacc_hip_histogram": (c_int32, [POINTER(View), c_int32, c_float, c_void_p, c_void_p]),
    "hipacc_hip_bilateral": (c_int32, [POINTER(BilateralDesc), c_void_p]),
    "hipacc_hip_local_grad_workspace_bytes": (c_size_t, [POINTER(LocalGradDesc)]),
    "hipacc_hip_bilateral_grad_workspace_bytes": (c_size_t, [POINTER(BilateralGradDesc)]),
    "hipacc_hip_resample": (c_int32, [POINTER(ResampleDesc), c_void_p]),
    "hipacc_hip_pyramid_plan_create": (c_int32, [POINTER(PyramidDesc), c_void_p,
                                                 POINTER(c_void_p)]),
    "hipacc_hip_pyramid_run": (c_int32, [c_void_p, c_void_p]),
    "hipacc_hip_pyramid_run_direct": (c_int32, [c_void_p, c_void_p]),
    "hipacc_hip_pyramid_plan_destroy": (c_int32, [c_void_p]),
    "hipacc_hip_pyramid_plan_tiles": (c_int32, [c_void_p, c_void_p, c_int32, c_int32, c_int32,
                                                c_void_p, c_void_p]),
    "hipacc_hip_local_grad": (c_int32, [POINTER(LocalGradDesc), c_void_p]),
    "hipacc_hip_local_grad_variant": (c_int32, [POINTER(LocalGradDesc), c_int32, c_void_p]),
    "hipacc_hip_local_grad_noalloc": (c_int32, [POINTER(LocalGradDesc), c_void_p]),
    "hipacc_hip_bilateral_train": (c_int32, [POINTER(BilateralGradDesc), c_void_p]),
    "hipacc_hip_bilateral_grad": (c_int32, [POINTER(BilateralGradDesc), c_void_p]),
    "hipacc_hip_alloc2d": (c_int32, [POINTER(c_void_p), POINTER(c_int32), c_int32, c_int32, c_int32,
                                     c_int32]),
    "hipacc_hip_free": (c_int32, [c_void_p]),
    "hipacc_hip_upload2d": (c_int32, [c_void_p, c_int32, c_void_p, c_int32, c_int32, c_int32,
                                      c_void_p]),
    "hipacc_hip_download2d": (c_int32, [c_void_p, c_void_p, c_int32, c_int32, c_int32, c_int32,
                                        c_void_p]),
    "hipacc_hip_copy_region": (c_int32, [POINTER(View), POINTER(View), c_void_p]),
    "hipacc_hip_fill": (c_int32, [c_void_p, c_uint32, c_size_t, c_int32, c_void_p]),
    "hipacc_hip_fill_random": (c_int32, [POINTER(View), c_uint64, c_void_p]),
    "hipacc_hip_set_timing": (c_int32, [c_int32]),
    "hipacc_hip_last_kernel_timing": (c_float, []),
    "hipacc_hip_get_timing": (c_int32, []),
    "hipacc_hip_set_last_kernel_timing": (c_int32, [c_float]),
    "hipacc_hip_local_variant": (c_int32, [POINTER(LocalDesc), c_int32, c_int32, c_void_p]),
    "hipacc_hip_separable_variant": (c_int32, [POINTER(SeparableDesc), c_int32, c_int32, c_void_p]),
    "hipacc_hip_tune_local": (c_int32, [POINTER(LocalDesc), c_int32, POINTER(TuneEntry), c_void_p]),
    "hipacc_hip_tune_separable": (c_int32, [POINTER(SeparableDesc), c_int32, POINTER(TuneEntry),
                                            c_void_p]),
    "hipacc_hip_tune_set": (c_int32, [POINTER(TuneEntry)]),
    "hipacc_hip_tune_save": (c_int32, [ctypes.c_char_p]),
    "hipacc_hip_tune_load": (c_int32, [ctypes.c_char_p]),
}

_LIB = None


class HipaccError(RuntimeError):
    """A non-zero status returned by the C ABI."""

    def __init__(self, status: int, where: str, detail: str):
        super().__init__(f"{where}: status {status} ({_status_name(status)}): {detail}")
        self.status = status


def _status_name(st: int) -> str:
    return {0: "ok", 1: "invalid argument", 2: "unsupported", 3: "HIP runtime error",
            4: "no usable device", 5: "out of device memory"}.get(st, "unknown")


def lib() -> ctypes.CDLL:
    """Loads libhipacc_gfx950.so; raises when it has not been built."""
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m hipacc_amd.build`")
        L = ctypes.CDLL(str(LIB_PATH), mode=os.RTLD_LOCAL if hasattr(os, "RTLD_LOCAL") else 0)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def check(status: int, where: str) -> None:
    if status != OK:
        raise HipaccError(status, where, lib().hipacc_hip_last_error().decode())
