"""The C++ runtime (hipacc_amd/runtime/hipacc_hip.hpp) driven by a host
program, tests/cpp/runtime_sample.cpp, that uses it the way generated host
code uses the reference runtime (runtime/hipacc_cu.hpp:86-372,
runtime/hipacc_cu.tpp:33-497, runtime/hipacc_base.hpp:131-191).  The program
runs as a child process on the GPU; its outputs are checked here against the
oracle on the same inputs: bit-exact for integer results and for the float
local / separable / point results, the float sum within rounding."""
from __future__ import annotations

import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle_binding as ob
from hipacc_amd import capi

ROOT = Path(__file__).resolve().parents[1]
SAMPLE = ROOT / "tests" / "cpp" / "runtime_sample"
W, H = 203, 117

G5 = np.array([0.00296902, 0.01330621, 0.02193823, 0.01330621, 0.00296902,
               0.01330621, 0.05963430, 0.09832033, 0.05963430, 0.01330621,
               0.02193823, 0.09832033, 0.16210282, 0.09832033, 0.02193823,
               0.01330621, 0.05963430, 0.09832033, 0.05963430, 0.01330621,
               0.00296902, 0.01330621, 0.02193823, 0.01330621, 0.00296902], np.float32)
K7 = np.array([0.0366328, 0.111281, 0.216745, 0.270682, 0.216745, 0.111281, 0.0366328],
              np.float32)


def _inputs():
    y, x = np.mgrid[0:H, 0:W]
    v = ((x * 37 + y * 91 + (x * y) % 13) & 255).astype(np.uint32)
    return v.astype(np.uint8), (v.astype(np.float32) / np.float32(255.0)).astype(np.float32)


@pytest.fixture(scope="module")
def results(tmp_path_factory):
    if not SAMPLE.exists():
        pytest.fail(f"{SAMPLE} is not built (make -C tests/cpp)")
    out = tmp_path_factory.mktemp("runtime_sample")
    print(f"running {SAMPLE} {out}", flush=True)
    p = subprocess.run([str(SAMPLE), str(out)], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, f"runtime_sample failed ({p.returncode}): {p.stderr}"
    scal = {}
    for line in (out / "scalars.txt").read_text().splitlines():
        k, *vals = line.split()
        scal[k] = vals
    return out, scal


def _load(out, name, dtype, shape):
    return np.fromfile(out / f"{name}.bin", dtype=dtype).reshape(shape)


@pytest.mark.gpu
def test_runtime_local_and_separable(results):
    out, scal = results
    assert scal["status"] == ["0"]
    u8, f32 = _inputs()
    ref = np.zeros_like(u8)
    ob.local(u8, ref, 5, 5, weights=G5, boundary=capi.CLAMP, cvt=capi.CVT_ROUND)
    assert np.array_equal(_load(out, "gauss_u8", np.uint8, (H, W)), ref)

    sx = np.array([-1, 0, 1, -2, 0, 2, -1, 0, 1], np.int32)
    ref16 = np.zeros((H, W), np.int16)
    ob.local(u8, ref16, 3, 3, weights=sx, domain=(sx != 0).astype(np.uint8), acc_float=False,
             boundary=capi.MIRROR)
    assert np.array_equal(_load(out, "sobel_i16", np.int16, (H, W)), ref16)

    sep = np.zeros_like(f32)
    ob.separable(f32, sep, K7, K7, boundary=capi.REPEAT)
    assert np.array_equal(_load(out, "sep_f32", np.float32, (H, W)), sep)

    diff = np.zeros_like(f32)
    ob.point(capi.POINT_ABSDIFF, f32, diff, sep)
    assert np.array_equal(_load(out, "absdiff_f32", np.float32, (H, W)), diff)


@pytest.mark.gpu
def test_runtime_windows_and_copies(results):
    out, scal = results
    u8, f32 = _inputs()
    win = np.zeros((60, 120), np.float32)
    ob.local(f32, win, 3, 3, reduce=capi.MAX, boundary=capi.CONSTANT,
             src_window=(5, 7, 100, 50), dst_window=(3, 2, 100, 50))
    assert np.array_equal(_load(out, "window_max_f32", np.float32, (60, 120)), win)

    cpy = np.zeros((32, 64), np.uint8)
    cpy[6:26, 4:44] = u8[13:33, 11:51]
    assert np.array_equal(_load(out, "copy_u8", np.uint8, (32, 64)), cpy)
    # 256-byte pitched rows for the aligned image, dense rows otherwise
    assert scal["stride_gauss"] == ["256"] and scal["stride_sobel"] == [str(W)]


@pytest.mark.gpu
def test_runtime_global_operators(results):
    out, scal = results
    u8, f32 = _inputs()
    assert np.float32(scal["fmin"][0]) == f32.min()
    assert np.float32(scal["fmax"][0]) == f32[30:70, 20:70].max()
    assert int(scal["usum"][0]) == int(u8.astype(np.uint64).sum())
    exact = float(np.sum(f32.astype(np.float64)))
    assert abs(float(scal["fsum"][0]) - exact) <= 1e-5 * exact
    assert np.array_equal(_load(out, "hist_u32", np.uint32, (256,)),
                          np.bincount(u8.ravel(), minlength=256).astype(np.uint32))
    # the generated-code names: hipaccApplyReductionShared / BinningSegmented
    assert scal["fmax_shared"] == scal["fmax"]
    assert np.array_equal(_load(out, "hist2_u32", np.uint32, (256,)),
                          np.bincount(u8.ravel(), minlength=256).astype(np.uint32))


@pytest.mark.gpu
def test_runtime_pyramid_traversal(results):
    out, scal = results
    _, f32 = _inputs()
    assert scal["visited"] == ["4"]
    w3, h3 = W >> 3, H >> 3
    assert scal["pyr3"] == [str(w3), str(h3)]
    assert np.array_equal(_load(out, "pyr3_f32", np.float32, (h3, w3)), f32[:h3, :w3])


@pytest.mark.gpu
def test_runtime_kernel_timing(results):
    """hipaccSetTiming(true): the launch is timed (a positive time);
    hipaccSetTiming(false): hipacc_last_kernel_timing keeps the last time."""
    out, scal = results
    t_on, t_off = (float(v) for v in scal["timing"])
    assert t_on > 0.0
    assert t_off == t_on


LAUNCH = ROOT / "tests" / "cpp" / "launch_sample"


@pytest.mark.gpu
def test_own_kernel_launched_through_the_runtime(gpu, tmp_path):
    """hipacc_launch_info / hipaccCalcGridFromBlock / hipaccPrepareKernelLaunch
    / hipaccLaunchKernel (runtime/hipacc_base.hpp:87-101,
    runtime/hipacc_cu_standalone.hpp:66-110,291-329) with a kernel of the
    program's own: outputs equal the host's, no block inside the border
    bounds reads outside the image, the timed launch reports a time."""
    assert LAUNCH.exists(), f"{LAUNCH} is not built (make -C tests/cpp)"
    print("launch_sample", flush=True)
    p = subprocess.run([str(LAUNCH)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(p.stdout)
    assert p.returncode == 0 and "PASSED" in p.stdout, p.stdout + p.stderr
