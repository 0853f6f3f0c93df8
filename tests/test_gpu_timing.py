"""hipacc_hip_set_timing / hipacc_hip_last_kernel_timing (the reference's
hipacc_last_kernel_timing, runtime/hipacc_base.hpp:54-66, whose state is
thread_local, runtime/hipacc_base_standalone.hpp:35-39): a timed launch
gives a positive time; a call with timing off, or a call that fails, leaves
it unchanged; another host thread has its own switch and last time."""
import ctypes
import threading

import numpy as np
import pytest
import torch

from hipacc_amd import capi, masks, operators as op

pytestmark = pytest.mark.gpu


def test_timing_on_off_failure_and_threads(gpu):
    L = capi.lib()
    src = torch.rand((512, 640), device="cuda")
    dst = torch.empty_like(src)
    call = op.prepare_local(src, dst, 5, 5, weights=masks.gaussian_2d(5, 1.0),
                            boundary=capi.CLAMP)
    try:
        assert L.hipacc_hip_set_timing(1) == 0
        call.launch()
        t1 = L.hipacc_hip_last_kernel_timing()
        assert t1 > 0.0
        # a failing call returns its status and leaves the time unchanged
        bad = op.prepare_local(src, dst, 5, 5, weights=masks.gaussian_2d(5, 1.0))
        bad.desc.mask_w = 4
        st = L.hipacc_hip_local(ctypes.byref(bad.desc),
                                torch.cuda.current_stream().cuda_stream)
        assert st == 1  # HIPACC_ERR_INVALID_ARG
        assert L.hipacc_hip_last_kernel_timing() == t1
        # timing off: launches are not timed
        assert L.hipacc_hip_set_timing(0) == 0
        call.launch()
        torch.cuda.synchronize()
        assert L.hipacc_hip_last_kernel_timing() == t1
        # another host thread: its own switch (off) and last time (0)
        seen = []

        def worker():
            seen.append(L.hipacc_hip_last_kernel_timing())
            call.launch()
            torch.cuda.synchronize()
            seen.append(L.hipacc_hip_last_kernel_timing())
            L.hipacc_hip_set_timing(1)
            call.launch()
            seen.append(L.hipacc_hip_last_kernel_timing())
            L.hipacc_hip_set_timing(0)

        th = threading.Thread(target=worker)
        th.start()
        th.join()
        assert seen[0] == 0.0 and seen[1] == 0.0 and seen[2] > 0.0, seen
        assert L.hipacc_hip_last_kernel_timing() == t1   # this thread's value
        # the switch as the runtime reads it, and a time recorded for a
        # launch made outside the library (a negative one is refused)
        assert L.hipacc_hip_get_timing() == 0
        L.hipacc_hip_set_timing(1)
        assert L.hipacc_hip_get_timing() == 1
        assert L.hipacc_hip_set_last_kernel_timing(ctypes.c_float(2.5)) == 0
        assert L.hipacc_hip_last_kernel_timing() == 2.5
        assert L.hipacc_hip_set_last_kernel_timing(ctypes.c_float(-1.0)) == 1
        assert L.hipacc_hip_last_kernel_timing() == 2.5
    finally:
        L.hipacc_hip_set_timing(0)
