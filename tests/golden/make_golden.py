"""Writes tests/golden/dsl_golden.npz: outputs of the reference's own DSL host
emulation (oracle/_ref/ref_dsl, built by `make -C oracle ref` from the
reference checkout's dsl/ headers) for small seeded inputs.

Each case is (op, width, height, seed, boundary, args); the inputs are not
stored -- they are SplitMix64 pixels regenerated from the seed (see
oracle/splitmix.h).  Run from the repository root:

    make -C oracle ref && python tests/golden/make_golden.py

UNDEFINED is not generated: the DSL emulation reads outside the image for it.
CONSTANT uses the value 0 (the DSL ignores the user's constant and uses
data_t{}, dsl/image.hpp:260-291).
"""
from __future__ import annotations

import hashlib
import json
import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
REF = ROOT / "oracle" / "_ref" / "ref_dsl"

MODES = {"CLAMP": 1, "REPEAT": 2, "MIRROR": 3, "CONSTANT": 4}
SIZES = [(37, 29), (64, 48), (130, 67)]


def cases():
    seed = 1000
    out = []

    def add(op, w, h, mode, args, dtype, shape=None):
        nonlocal seed
        seed += 1
        out.append(dict(op=op, w=w, h=h, seed=seed, boundary=MODES[mode], mode=mode,
                        args=[str(a) for a in args], dtype=dtype,
                        shape=shape if shape is not None else [h, w]))

    for (w, h) in SIZES:
        for m in MODES:
            add("conv_u8", w, h, m, [5, 1.0], "uint8")
            add("sobel_u8", w, h, m, [0], "int16")
            add("sobel_u8", w, h, m, [1], "int16")
            add("box_u8", w, h, m, [7], "uint32")
            add("min_u8", w, h, m, [3], "uint8")
    for (w, h) in SIZES[:2]:
        for m in MODES:
            add("conv_f32", w, h, m, [5, 1.0], "float32")
            add("conv_f32", w, h, m, [3, 0.8], "float32")
            add("max_f32", w, h, m, [3], "float32")
            add("conv_u8", w, h, m, [13, 3.0], "uint8")
    for m in MODES:
        add("conv_f32", 37, 29, m, [13, 3.0], "float32")
        add("conv_f32", 37, 29, m, [9, 2.0], "float32")
    for m in ("MIRROR", "CLAMP", "REPEAT", "CONSTANT"):
        add("bilat_f32", 37, 29, m, [13, 3.0, 0.1], "float32")
        add("bilat_f32", 64, 48, m, [7, 2.0, 0.2], "float32")
    for (w, h) in SIZES:
        for op in (0, 1, 2):
            add("red_f32", w, h, "CLAMP", [op], "float32", [1])
        for op in (1, 2):
            add("red_u8", w, h, "CLAMP", [op], "uint8", [1])
        add("hist_u8", w, h, "CLAMP", [256], "uint32", [256])
        add("hist_u8", w, h, "CLAMP", [64], "uint32", [64])
        add("sub_i8", w, h, "CLAMP", [], "int8")
    # resampling accessors used by pyramids: odd and even source sizes
    for (w, h, dw, dh) in ((37, 29, 18, 14), (64, 48, 32, 24), (18, 14, 37, 29), (32, 24, 64, 48)):
        add("nn_f32", w, h, "CLAMP", [dw, dh], "float32", [dh, dw])
        add("lf_f32", w, h, "CLAMP", [dw, dh], "float32", [dh, dw])
    # the whole interpolation set (NN, LF, B5, CF, L3) on float and uchar
    # accessors: down / up sampling, a cropped accessor window with call
    # offsets in(xf, yf), each boundary mode (REPEAT on a whole image: on a
    # crop the DSL's single shift by lower + upper differs from the generated
    # code's periodic wrap, which the library follows)
    geoms = (  # w, h, dst_w, dst_h, xf, yf, crop (x, y, w, h; 0 = whole), mode
        (37, 29, 18, 14, 0, 0, (0, 0, 0, 0), "CLAMP"),
        (18, 14, 37, 29, 0, 0, (0, 0, 0, 0), "MIRROR"),
        (40, 32, 30, 25, 1, -2, (5, 3, 20, 17), "CLAMP"),
        (29, 35, 45, 17, 2, 1, (0, 0, 0, 0), "REPEAT"),
        (33, 21, 50, 40, -1, 1, (0, 0, 0, 0), "CONSTANT"),
    )
    # 4-wide vector pixels: the RGBA samples' operators
    for (w, h) in SIZES[:2]:
        for m in MODES:
            add("conv_rgba", w, h, m, [5, 1.0], "uint8", [h, w, 4])
            add("conv_rgba", w, h, m, [3, 0.8], "uint8", [h, w, 4])
            add("conv_f32x4", w, h, m, [5, 1.0], "float32", [h, w, 4])
            add("box_rgba", w, h, m, [5, 0], "uint8", [h, w, 4])
            add("erode_rgba", w, h, m, [3, 0], "uint8", [h, w, 4])
            add("sobel_rgba", w, h, m, [1], "int32", [h, w, 4])
    for imode in (1, 2, 3, 4, 5):
        for dt, op in (("float32", "interp_f32"), ("uint8", "interp_u8")):
            for (w, h, dw, dh, xf, yf, crop, m) in geoms:
                add(op, w, h, m, [imode, dw, dh, xf, yf, *crop], dt, [dh, dw])
    # the kernel chains of the fusion samples (P2P, P2L, L2P, L2L, Mixed):
    # the unfused chain's output, which a fused launch must equal
    for (w, h) in SIZES[:2]:
        for m in MODES:
            for st in ("PPP", "PPL", "LPP", "LL", "PPLPPLPP", "PLPL"):
                if m == "CLAMP" or "L" in st:
                    add("fuse_u8", w, h, m, [st], "uint8")
            add("fuse_f32", w, h, m, ["PLPLP"], "float32")
    return out


def main() -> int:
    if not REF.exists():
        print(f"{REF} missing: run `make -C oracle ref` first", file=sys.stderr)
        return 1
    arrays = {}
    manifest = []
    with tempfile.TemporaryDirectory() as td:
        for i, c in enumerate(cases()):
            path = Path(td) / f"{i}.bin"
            cmd = [str(REF), c["op"], str(c["w"]), str(c["h"]), str(c["seed"]), str(c["boundary"]),
                   str(path), *c["args"]]
            subprocess.run(cmd, check=True)
            data = np.fromfile(path, dtype=np.dtype(c["dtype"])).reshape(c["shape"])
            key = f"case{i:03d}"
            arrays[key] = data
            c["key"] = key
            c["sha256"] = hashlib.sha256(data.tobytes()).hexdigest()
            manifest.append(c)
    np.savez_compressed(HERE / "dsl_golden.npz", **arrays)
    (HERE / "dsl_golden.json").write_text(json.dumps(manifest, indent=1) + "\n")
    print(f"wrote {len(manifest)} cases")
    return 0


if __name__ == "__main__":
    sys.exit(main())
