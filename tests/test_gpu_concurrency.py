"""Host threads sharing one GPU through the C ABI: blocking reductions
(reduced_data()) from two threads on two streams return each thread's own
result, and histograms on two streams use their own per-stream workspaces."""
import ctypes
import threading

import numpy as np
import pytest
import torch

from hipacc_amd import capi, operators as op

pytestmark = pytest.mark.gpu


def _run_threads(fns):
    errors = []

    def wrap(fn):
        try:
            fn()
        except Exception as e:  # noqa: BLE001 - reported below
            errors.append(repr(e))

    ts = [threading.Thread(target=wrap, args=(f,)) for f in fns]
    for t in ts:
        t.start()
    for t in ts:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in ts), "a thread did not finish"
    assert not errors, errors


def test_reduce_sync_from_two_threads(gpu):
    L = capi.lib()
    imgs = [torch.full((1024, 1024), 1.0, device="cuda"),
            torch.full((1024, 1024), 2.0, device="cuda")]
    expected = [1024.0 * 1024.0, 2.0 * 1024.0 * 1024.0]   # exact in float32
    torch.cuda.synchronize()
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]

    def worker(k):
        def run():
            v = op.view(imgs[k])
            out = np.zeros(1, np.float32)
            for i in range(200):
                st = L.hipacc_hip_reduce_sync(ctypes.byref(v), capi.SUM, capi.F32,
                                              out.ctypes.data, streams[k].cuda_stream)
                capi.check(st, "hipacc_hip_reduce_sync")
                if out[0] != expected[k]:
                    raise AssertionError(f"thread {k} call {i}: {out[0]} != {expected[k]}")
        return run

    _run_threads([worker(0), worker(1)])


def test_histograms_on_two_streams(gpu):
    L = capi.lib()
    imgs = [torch.full((2048, 2048), 7, dtype=torch.uint8, device="cuda"),
            torch.full((2048, 2048), 200, dtype=torch.uint8, device="cuda")]
    bins = [torch.zeros(256, dtype=torch.int32, device="cuda") for _ in range(2)]
    torch.cuda.synchronize()
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]

    def worker(k):
        def run():
            v = op.view(imgs[k])
            for _ in range(100):
                st = L.hipacc_hip_histogram(ctypes.byref(v), 256, 1.0, bins[k].data_ptr(),
                                            streams[k].cuda_stream)
                capi.check(st, "hipacc_hip_histogram")
            streams[k].synchronize()
        return run

    _run_threads([worker(0), worker(1)])
    torch.cuda.synchronize()
    for k, value in enumerate((7, 200)):
        h = bins[k].cpu().numpy()
        assert h[value] == 2048 * 2048 and h.sum() == 2048 * 2048
