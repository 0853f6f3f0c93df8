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
    if op in ("conv_u8", "sobel_u8", "box_u8", "min_u8", "red_u8", "hist_u8", "interp_u8"):
        return ob.random_image((h, w), np.uint8, seed)
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
