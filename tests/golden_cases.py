"""Loads the reference-DSL golden vectors and maps each case onto the
descriptor the oracle and the GPU path take."""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np

from hipacc_amd import capi, masks

import oracle_binding as ob

HERE = Path(__file__).resolve().parent / "golden"


def load():
    manifest = json.loads((HERE / "dsl_golden.json").read_text())
    data = np.load(HERE / "dsl_golden.npz", allow_pickle=False)
    return [(c, data[c["key"]]) for c in manifest]


def case_id(c) -> str:
    return f'{c["op"]}-{"x".join(c["args"]) or "na"}-{c["w"]}x{c["h"]}-{c["mode"]}'


def input_for(c):
    """The SplitMix64 input(s) the reference driver generated for a case."""
    w, h, seed = c["w"], c["h"], c["seed"]
    op = c["op"]
    if op in ("conv_u8", "sobel_u8", "box_u8", "min_u8", "red_u8", "hist_u8", "interp_u8",
              "fuse_u8"):
        return ob.random_image((h, w), np.uint8, seed)
    if op in ("conv_rgba", "box_rgba", "erode_rgba", "sobel_rgba"):
        return ob.random_image((h, 4 * w), np.uint8, seed).reshape(h, w, 4)
    if op == "conv_f32x4":
        return ob.random_image((h, 4 * w), np.float32, seed).reshape(h, w, 4)
    if op == "sub_i8":
        return (ob.random_image((h, w), np.int8, seed), ob.random_image((h, w), np.int8, seed + 1))
    return ob.random_image((h, w), np.float32, seed)


def local_spec(c):
    """kwargs of a local-operator call reproducing the case (or None)."""
    op, a = c["op"], c["args"]
    b = c["boundary"]
    if op in ("conv_u8", "conv_f32"):
        k, s = int(a[0]), float(a[1])
        return dict(mask_w=k, mask_h=k, weights=masks.gaussian_2d(k, s), acc_float=True,
                    boundary=b, cvt=capi.CVT_ROUND if op == "conv_u8" else capi.CVT_TRUNC,
                    fp_contract=False,
                    out_dtype=np.uint8 if op == "conv_u8" else np.float32)
    if op == "sobel_u8":
        m = masks.SOBEL_Y if int(a[0]) else masks.SOBEL_X
        return dict(mask_w=3, mask_h=3, weights=m, domain=(m != 0).astype(np.uint8),
                    acc_float=False, boundary=b, out_dtype=np.int16)
    if op == "box_u8":
        k = int(a[0])
        return dict(mask_w=k, mask_h=k, acc_float=False, boundary=b, out_dtype=np.uint32)
    if op == "min_u8":
        k = int(a[0])
        return dict(mask_w=k, mask_h=k, acc_float=False, reduce=capi.MIN, boundary=b,
                    out_dtype=np.uint8)
    if op == "max_f32":
        k = int(a[0])
        return dict(mask_w=k, mask_h=k, acc_float=True, reduce=capi.MAX, boundary=b,
                    out_dtype=np.float32)
    # 4-wide vector pixels (the RGBA samples)
    if op == "conv_rgba":
        k, s = int(a[0]), float(a[1])
        return dict(mask_w=k, mask_h=k, weights=masks.gaussian_2d(k, s), acc_float=True,
                    boundary=b, cvt=capi.CVT_ROUND, fp_contract=False, src_dtype=capi.U8X4,
                    dst_dtype=capi.U8X4, out_dtype=np.uint8)
    if op == "conv_f32x4":
        k, s = int(a[0]), float(a[1])
        return dict(mask_w=k, mask_h=k, weights=masks.gaussian_2d(k, s), acc_float=True,
                    boundary=b, fp_contract=False, src_dtype=capi.F32X4, dst_dtype=capi.F32X4,
                    out_dtype=np.float32)
    if op == "box_rgba":
        k = int(a[0])
        return dict(mask_w=k, mask_h=k, acc_float=False, boundary=b, post_div=float(k * k),
                    src_dtype=capi.U8X4, dst_dtype=capi.U8X4, out_dtype=np.uint8)
    if op == "erode_rgba":
        k = int(a[0])
        return dict(mask_w=k, mask_h=k, acc_float=False, reduce=capi.MIN, boundary=b,
                    src_dtype=capi.U8X4, dst_dtype=capi.U8X4, out_dtype=np.uint8)
    if op == "sobel_rgba":
        m = masks.SOBEL_Y if int(a[0]) else masks.SOBEL_X
        return dict(mask_w=3, mask_h=3, weights=m, domain=(m != 0).astype(np.uint8),
                    acc_float=False, boundary=b, src_dtype=capi.U8X4, dst_dtype=capi.I32X4,
                    out_dtype=np.int32)
    return None


def interp_spec(c):
    """(interp, dst shape, kwargs) of a resample call reproducing an
    interp_f32 / interp_u8 case."""
    a = [int(v) for v in c["args"]]
    imode, dw, dh, xf, yf, cx, cy, cw, ch = a
    kw = dict(boundary=c["boundary"], xf=xf, yf=yf)
    if cw > 0:
        kw["src_window"] = (cx, cy, cw, ch)
    return imode, (dh, dw), kw


# The 5x5 mask of the fusion samples (samples-public/6_Test/Kernel_Fusion_P2L/
# src/main.cpp:91-97).
FUSE_MASK = np.array([[0.005008, 0.017300, 0.026151, 0.017300, 0.005008],
                      [0.017300, 0.059761, 0.090339, 0.059761, 0.017300],
                      [0.026151, 0.090339, 0.136565, 0.090339, 0.026151],
                      [0.017300, 0.059761, 0.090339, 0.059761, 0.017300],
                      [0.005008, 0.017300, 0.026151, 0.017300, 0.005008]], np.float32)


def fuse_spec(c):
    """The stage list (operators.prepare_fused / oracle_binding.chain form)
    of a fuse_u8 / fuse_f32 case: P = AFFINE v + 3 (the sample's `v += 3`),
    L = the 5x5 mask, SUM, + 0.5 rounding, without contraction (the DSL
    host emulation)."""
    dt = capi.U8 if c["op"] == "fuse_u8" else capi.F32
    stages = []
    for ch in c["args"][0]:
        if ch == "P":
            stages.append(dict(kind="point", dtype=dt, op=capi.POINT_AFFINE, alpha=1.0, beta=3.0))
        else:
            stages.append(dict(kind="local", dtype=dt, mask_w=5, mask_h=5,
                               weights=FUSE_MASK.reshape(-1), boundary=c["boundary"],
                               cvt=capi.CVT_ROUND, fp_contract=False))
    return stages
