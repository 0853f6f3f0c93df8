"""hipacc_hip_local_grad_noalloc, the allocation-free adjoint: through the C
ABI, against hipacc_hip_local_grad on the same inputs (same kernels, so the
same bits), on two calls in a row with the second smaller, and its refusal
of descriptors whose workspace exceeds the static scratch array."""
import ctypes

import numpy as np
import pytest
import torch

from hipacc_amd import capi, operators as op

pytestmark = pytest.mark.gpu


def _inputs(shape, k, seed):
    g = torch.Generator(device="cuda").manual_seed(seed)
    x = torch.rand(shape, generator=g, device="cuda")
    gy = torch.randn(shape, generator=g, device="cuda")
    w = (0.1 * torch.randn(k * k, generator=g, device="cuda")).contiguous()
    return x, gy, w


def test_noalloc_equals_local_grad(gpu):
    L = capi.lib()
    stream = torch.cuda.current_stream().cuda_stream
    # the second call is smaller than the first: the scratch array keeps the
    # first call's contents, which must not leak into the second result
    for shape, k, mode in (((4, 640, 768), 9, capi.CLAMP), ((2, 301, 257), 5, capi.MIRROR)):
        print(f"noalloc adjoint shape={shape} k={k} mode={mode}", flush=True)
        x, gy, w = _inputs(shape, k, 11 + k)
        gx_ref, gw_ref = torch.empty_like(x), torch.empty(k * k, device="cuda")
        op.prepare_local_grad(x, gy, gx_ref, k, k, weights_dev=w, grad_weights=gw_ref,
                              boundary=mode).launch()
        gx, gw = torch.full_like(x, -7.0), torch.full((k * k,), -7.0, device="cuda")
        call = op.prepare_local_grad(x, gy, gx, k, k, weights_dev=w, grad_weights=gw,
                                     boundary=mode)
        assert L.hipacc_hip_local_grad_workspace_bytes(ctypes.byref(call.desc)) <= \
            capi.NOALLOC_SCRATCH_BYTES
        st = L.hipacc_hip_local_grad_noalloc(ctypes.byref(call.desc), stream)
        capi.check(st, "hipacc_hip_local_grad_noalloc")
        torch.cuda.synchronize()
        assert torch.equal(gx, gx_ref), "dX differs from hipacc_hip_local_grad"
        assert torch.equal(gw, gw_ref), "dw differs from hipacc_hip_local_grad"


def test_noalloc_refuses_oversized_and_workspace(gpu):
    L = capi.lib()
    stream = torch.cuda.current_stream().cuda_stream
    x, gy, w = _inputs((1, 64, 64), 15, 3)
    gx, gw = torch.empty_like(x), torch.empty(225, device="cuda")
    call = op.prepare_local_grad(x, gy, gx, 15, 15, weights_dev=w, grad_weights=gw)
    d = call.desc
    # a 64 x 4096 x 4096 batch: its halo ring alone is ~29 MB; refused on the
    # host before any launch (the pointers are never dereferenced)
    for v in (d.input, d.grad_out, d.grad_in):
        v.width = v.height = v.img_width = v.img_height = v.stride = 4096
        v.batch, v.plane_stride = 64, 4096 * 4096
    assert L.hipacc_hip_local_grad_workspace_bytes(ctypes.byref(d)) > capi.NOALLOC_SCRATCH_BYTES
    st = L.hipacc_hip_local_grad_noalloc(ctypes.byref(d), stream)
    assert st == 1 and b"static scratch" in L.hipacc_hip_last_error()
    # a caller-provided workspace is refused (the scratch array is the workspace)
    ws = torch.empty(1 << 20, dtype=torch.uint8, device="cuda")
    call2 = op.prepare_local_grad(x, gy, gx, 15, 15, weights_dev=w, grad_weights=gw, workspace=ws)
    st = L.hipacc_hip_local_grad_noalloc(ctypes.byref(call2.desc), stream)
    assert st == 1 and b"workspace must be NULL" in L.hipacc_hip_last_error()
    torch.cuda.synchronize()
