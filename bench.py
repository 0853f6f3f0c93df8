"""Benchmark of the data-parallel operator path on one MI355X.

Headline (``value``): 4096x4096 float32 5x5 Gaussian (sigma 1) convolution
with CLAMP boundary handling, in Mpixel/s, with its HBM roofline fraction at
8 B/pixel algorithmic traffic (read 4 B + write 4 B).  A "step" is one
operator launch over one 4096x4096 image.  Four input/output buffer pairs
(512 MiB) are rotated so the working set exceeds the 256 MB Infinity Cache.

The same run also times the rest of the 4096x4096 operator suite (Sobel,
separable 13x13 Gaussian, 7x7 box sum, reductions, histogram, bilateral) and
reports them under ``ops``; the CPU baseline is the C oracle (OpenMP
restatement of the reference's generated CPU backend) on a bounded sample.

Before timing, the headline operator's kernel variant is chosen by timing
each candidate (row-streaming segment heights, wide streaming, LDS tile) on
the same rotating buffers; the choice is reported under roofline.variant.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--quick]
    python bench.py --profile [--variant stream:12]   # kernels for rocprofv3
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import statistics
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
W = H = 4096
NPIX = W * H
NSETS = 4


def _events():
    import torch
    return torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)


def time_calls(calls, steps, warmup):
    """Launches calls[i % len(calls)] back to back; returns ms per launch
    measured with events on the launch stream (torch's current stream)."""
    import torch
    for i in range(warmup):
        calls[i % len(calls)].launch()
    torch.cuda.synchronize()
    e0, e1 = _events()
    e0.record()
    for i in range(steps):
        calls[i % len(calls)].launch()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


HEAD_CANDIDATES = ("stream:8", "stream:12", "stream:16", "stream:24", "wide:16", "tile:4")


def parse_variant(text):
    from hipacc_amd import capi
    name, _, param = text.partition(":")
    codes = {v: k for k, v in capi.VARIANT_NAMES.items()}
    return codes[name], int(param or 0)


def choose_variant(calls, steps=60):
    """Times each candidate variant on the rotating buffers; returns
    (ms, name, variant calls) of the fastest."""
    from hipacc_amd import capi
    best = None
    for name in HEAD_CANDIDATES:
        v, p = parse_variant(name)
        vc = [c.with_variant(v, p) for c in calls]
        try:
            ms = min(time_calls(vc, steps, 5) for _ in range(2))
        except capi.HipaccError:
            continue
        if best is None or ms < best[0]:
            best = (ms, name, vc)
    return best


def op_entry(name, ms, bytes_per_px, npix=NPIX, extra=None):
    gbs = bytes_per_px * npix / (ms * 1e-3) / 1e9
    d = {"ms": round(ms, 5), "mpix_s": round(npix / (ms * 1e-3) / 1e6, 1),
         "gbs": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 4), "bytes_per_px": bytes_per_px}
    if extra:
        d.update(extra)
    return d


def cpu_baseline(budget_s: float):
    """The C oracle (OpenMP over rows, per-pixel boundary handling) on whole
    4096x4096 images of the headline operator, repeated until `budget_s`."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_binding as ob
    from hipacc_amd import capi, masks
    src = ob.random_image((H, W), np.float32, 7)
    dst = np.zeros_like(src)
    w = masks.gaussian_2d(5, 1.0)
    ob.local(src, dst, 5, 5, weights=w, boundary=capi.CLAMP, fp_contract=True)  # warm-up
    n, t0 = 0, time.perf_counter()
    while True:
        ob.local(src, dst, 5, 5, weights=w, boundary=capi.CLAMP, fp_contract=True)
        n += 1
        el = time.perf_counter() - t0
        if el >= budget_s:
            break
    threads = ob.lib().oracle_num_threads()
    return {"value": round(n * NPIX / el / 1e6, 2), "unit": "Mpix/s", "cores": threads,
            "kind": "port",
            "sample": f"{n} full 4096x4096 float 5x5 CLAMP images ({el:.1f} s), "
                      f"oracle/hipacc_oracle.c with {threads} OpenMP threads"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--quick", action="store_true", help="headline operator only")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--profile", action="store_true",
                    help="launch the headline kernel 20 times and a 64 MiB copy 20 times, "
                         "for rocprofv3 counter runs; prints no result line")
    ap.add_argument("--variant", default="stream:12",
                    help="headline kernel variant under --profile (name:param)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the operator path runs on a single MI355X)", file=sys.stderr)
        return 2

    import torch
    from hipacc_amd import capi, masks, operators as op

    op.init(0)
    dev = torch.device("cuda:0")
    f32_in = [torch.empty((H, W), dtype=torch.float32, device=dev) for _ in range(NSETS)]
    f32_out = [torch.empty((H, W), dtype=torch.float32, device=dev) for _ in range(NSETS)]
    for i, t in enumerate(f32_in):
        op.fill_random(t, 1000 + i)
    torch.cuda.synchronize()

    # ---- headline: 5x5 float Gaussian, CLAMP ----------------------------------
    g5 = masks.gaussian_2d(5, 1.0)
    head = [op.prepare_local(f32_in[i], f32_out[i], 5, 5, weights=g5, boundary=capi.CLAMP,
                             fp_contract=True) for i in range(NSETS)]
    if args.profile:
        v, p = parse_variant(args.variant)
        calls = [c.with_variant(v, p) for c in head]
        for i in range(20):
            calls[i % NSETS].launch()
        src = torch.empty(16 << 20, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(20):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(f"profile run: 20 x {args.variant} 5x5 f32 4096^2, 20 x 64 MiB copy", flush=True)
        return 0
    _, vname, head = choose_variant(head)
    samples = [time_calls(head, args.steps, args.warmup if k == 0 else 2) for k in range(5)]
    ms = statistics.median(samples)
    e = op_entry("conv5x5_f32_clamp", ms, 8)
    result = {
        "metric": "Mpix/s + HBM GB/s (% peak), 4096^2 stencils; fwd+bwd vs plain fwd time",
        "value": e["mpix_s"],
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: SplitMix64 U[0,1) float pixels, generated on the GPU",
        "config": {"workload": "4096x4096 float32 5x5 Gaussian (sigma 1) convolution, CLAMP, "
                               "4 rotating buffer pairs (512 MiB)",
                   "image": [W, H], "mask": "5x5 gaussian sigma=1.0", "boundary": "CLAMP",
                   "fp_contract": True},
        "roofline": {"bound": "hbm", "achieved": e["gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": e["hbm_frac"], "traffic": None,
                     "kernel": "stream_kernel" if vname.startswith(("stream", "wide"))
                               else "local_sum_kernel",
                     "variant": vname,
                     "bytes_per_launch": 8 * NPIX,
                     "batch_ms": [round(s, 5) for s in samples]},
    }
    traffic = ROOT / "profiles" / "traffic.json"
    if traffic.exists():
        try:
            t = json.loads(traffic.read_text()).get("conv5x5_f32_clamp")
            if t:
                result["roofline"]["traffic"] = t
        except ValueError:
            pass

    ops = {"conv5x5_f32_clamp": e}
    if not args.quick:
        steps = max(20, args.steps // 2)
        u8_in = [torch.empty((H, W), dtype=torch.uint8, device=dev) for _ in range(NSETS)]
        u8_out = [torch.empty((H, W), dtype=torch.uint8, device=dev) for _ in range(NSETS)]
        i16_out = [torch.empty((H, W), dtype=torch.int16, device=dev) for _ in range(NSETS)]
        for i, t in enumerate(u8_in):
            op.fill_random(t, 2000 + i)
        torch.cuda.synchronize()
        sob = masks.SOBEL_X
        dom = (sob != 0).astype(np.uint8)
        g13 = masks.gaussian_1d(13, 3.0)
        ones7 = np.ones(7, np.int32)

        def run(name, calls, bpp, n=steps, extra=None, npix=NPIX):
            ops[name] = op_entry(name, time_calls(calls, n, 5), bpp, npix=npix, extra=extra)

        # 1024x1024 uint8 5x5 Gaussian, rounded uint8 output (cache-resident:
        # reported, no roofline claim)
        s_in = [torch.empty((1024, 1024), dtype=torch.uint8, device=dev) for _ in range(NSETS)]
        s_out = [torch.empty_like(t) for t in s_in]
        for i, t in enumerate(s_in):
            op.fill_random(t, 3000 + i)
        run("gauss5x5_u8_1024_clamp_round",
            [op.prepare_local(s_in[i], s_out[i], 5, 5, weights=g5, boundary=capi.CLAMP,
                              cvt=capi.CVT_ROUND) for i in range(NSETS)], 2, npix=1024 * 1024)

        run("sobel3x3_f32_clamp", [op.prepare_local(f32_in[i], f32_out[i], 3, 3, weights=sob,
                                                    boundary=capi.CLAMP) for i in range(NSETS)], 8)
        run("sobel3x3_u8_i16_clamp", [op.prepare_local(u8_in[i], i16_out[i], 3, 3, weights=sob,
                                                       domain=dom, acc_float=False,
                                                       boundary=capi.CLAMP)
                                      for i in range(NSETS)], 3)
        run("gauss13_sep_f32_clamp", [op.prepare_separable(f32_in[i], f32_out[i], g13, g13,
                                                           boundary=capi.CLAMP)
                                      for i in range(NSETS)], 8)
        run("gauss13_sep_u8_clamp", [op.prepare_separable(u8_in[i], u8_out[i], g13, g13,
                                                          boundary=capi.CLAMP, cvt=capi.CVT_ROUND)
                                     for i in range(NSETS)], 2)
        run("box7x7_u8_u32_clamp", [op.prepare_separable(u8_in[i], f32_out[i], ones7, ones7,
                                                         acc_float=False, boundary=capi.CLAMP,
                                                         dst_dtype=capi.U32)
                                    for i in range(NSETS)], 5)
        for mode, mname in ((capi.MIRROR, "mirror"), (capi.REPEAT, "repeat"),
                            (capi.CONSTANT, "constant"), (capi.UNDEFINED, "undefined")):
            run(f"conv5x5_f32_{mname}", [op.prepare_local(f32_in[i], f32_out[i], 5, 5, weights=g5,
                                                          boundary=mode) for i in range(NSETS)], 8)
        res_f = torch.zeros(4, dtype=torch.float32, device=dev)
        res_i = torch.zeros(4, dtype=torch.int32, device=dev)
        for mode, mname in ((capi.SUM, "sum"), (capi.MIN, "min"), (capi.MAX, "max")):
            run(f"reduce_{mname}_f32", [op.ReduceCall(f32_in[i], mode, res_f[i:i + 1])
                                        for i in range(NSETS)], 4)
        run("reduce_sum_u8_u32", [op.ReduceCall(u8_in[i], capi.SUM, res_i[i:i + 1])
                                  for i in range(NSETS)], 1)
        bins = torch.zeros((NSETS, 256), dtype=torch.int32, device=dev)
        run("histogram256_u8", [op.HistogramCall(u8_in[i], 256, bins[i]) for i in range(NSETS)], 1)
        bil = [op.prepare_bilateral(f32_in[i], f32_out[i], 13, 3.0, 0.1, boundary=capi.MIRROR)
               for i in range(NSETS)]
        bms = time_calls(bil, 12, 2)
        flop = NPIX * 169 * 8
        run_entry = op_entry("bilateral13_f32_mirror", bms, 8,
                             extra={"gflop_s": round(flop / (bms * 1e-3) / 1e9, 1),
                                    "flop_per_px": 169 * 8, "exp_per_px": 169})
        ops["bilateral13_f32_mirror"] = run_entry
        del u8_in, u8_out, i16_out
    result["ops"] = ops
    result["cpu_baseline"] = cpu_baseline(args.cpu_seconds if not args.quick else 2.0)
    print(json.dumps(result))
    return 0


if __name__ == "__main__":
    sys.exit(main())
