"""The C-ABI library without a GPU: it loads, exports every function the
header declares, its structs match the ctypes mirror, and descriptor errors
are reported as status codes (no device needed, nothing is launched)."""
import ctypes
import json
import re
import subprocess
from pathlib import Path

import pytest

from hipacc_amd import capi

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "hipacc_hip.h"


def declared_functions():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"\b(hipacc_hip_\w+)\s*\(", text)))


def test_library_exports_every_declared_function():
    L = capi.lib()
    names = declared_functions()
    assert names, "no declarations found"
    for n in names:
        assert hasattr(L, n), f"{n} declared in {HEADER.name} but not exported"
    assert set(names) == set(capi.SIGNATURES), set(names) ^ set(capi.SIGNATURES)


def test_struct_layout_matches_ctypes(tmp_path):
    exe = tmp_path / "abi_layout"
    subprocess.run(["gcc", "-std=c99", "-o", str(exe), str(ROOT / "tests" / "abi_layout.c")],
                   check=True)
    lay = json.loads(subprocess.run([str(exe)], check=True, capture_output=True,
                                    text=True).stdout)
    py = {"hipacc_view": capi.View, "hipacc_local_desc": capi.LocalDesc,
          "hipacc_separable_desc": capi.SeparableDesc, "hipacc_point_desc": capi.PointDesc,
          "hipacc_bilateral_desc": capi.BilateralDesc,
          "hipacc_local_grad_desc": capi.LocalGradDesc,
          "hipacc_bilateral_grad_desc": capi.BilateralGradDesc,
          "hipacc_tune_entry": capi.TuneEntry}
    rename = {"in": "input", "out": "output"}
    for key, off in lay.items():
        t, m = key.split(".")
        if t == "sizeof":
            cls = {"view": capi.View, "local": capi.LocalDesc, "sep": capi.SeparableDesc,
                   "point": capi.PointDesc, "bilat": capi.BilateralDesc,
                   "lgrad": capi.LocalGradDesc, "bgrad": capi.BilateralGradDesc,
                   "tune": capi.TuneEntry}[m]
            assert ctypes.sizeof(cls) == off, key
        else:
            assert getattr(py[t], rename.get(m, m)).offset == off, key


def test_version_and_strerror():
    L = capi.lib()
    assert L.hipacc_hip_version().startswith(b"hipacc-gfx950")
    assert L.hipacc_hip_strerror(0) == b"ok"
    assert L.hipacc_hip_strerror(2) == b"unsupported"
    assert L.hipacc_hip_strerror(99) == b"unknown status"


def _fake_view(w=64, h=32, dtype=capi.F32):
    v = capi.View()
    v.data = 0x10000  # never dereferenced: validation fails first
    v.dtype, v.stride, v.img_width, v.img_height = dtype, w, w, h
    v.width, v.height = w, h
    return v


@pytest.mark.parametrize("mutate,status", [
    (lambda d: setattr(d, "mask_w", 4), capi.HipaccError),
    (lambda d: setattr(d, "mask_w", 17), capi.HipaccError),
    (lambda d: setattr(d, "boundary", 7), capi.HipaccError),
    (lambda d: setattr(d, "reduce", 4), capi.HipaccError),
    (lambda d: setattr(d.input, "offset_x", 10), capi.HipaccError),
    (lambda d: setattr(d.output, "dtype", 9), capi.HipaccError),
    (lambda d: setattr(d, "acc_dtype", capi.U8), capi.HipaccError),
    (lambda d: setattr(d.input, "data", None), capi.HipaccError),
])
def test_local_descriptor_errors(mutate, status):
    d = capi.LocalDesc()
    d.input, d.output = _fake_view(), _fake_view()
    d.mask_w = d.mask_h = 3
    d.acc_dtype = capi.F32
    d.boundary = capi.CLAMP
    mutate(d)
    st = capi.lib().hipacc_hip_local(ctypes.byref(d), None)
    assert st == 1, capi.lib().hipacc_hip_last_error()
    assert capi.lib().hipacc_hip_last_error()


def test_other_entry_points_validate_before_launch():
    L = capi.lib()
    assert L.hipacc_hip_local(None, None) == 1
    assert L.hipacc_hip_separable(None, None) == 1
    assert L.hipacc_hip_point(None, None) == 1
    assert L.hipacc_hip_bilateral(None, None) == 1
    v = _fake_view()
    assert L.hipacc_hip_reduce(ctypes.byref(v), 9, capi.F32, ctypes.c_void_p(0x10), None, None) == 1
    assert L.hipacc_hip_histogram(ctypes.byref(v), 0, 1.0, ctypes.c_void_p(0x10), None) == 1
    s = capi.SeparableDesc()
    s.input, s.output = _fake_view(), _fake_view(w=32)
    s.ksize_x = s.ksize_y = 3
    s.acc_dtype = capi.F32
    assert L.hipacc_hip_separable(ctypes.byref(s), None) == 2  # IS != window: unsupported
    b = capi.BilateralDesc()
    b.input, b.output = _fake_view(), _fake_view()
    b.ksize, b.sigma_d, b.sigma_r = 5, 1.0, 0.0
    assert L.hipacc_hip_bilateral(ctypes.byref(b), None) == 1


def test_init_without_device_reports_no_device():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    assert capi.lib().hipacc_hip_init(0) == 4


def test_cpp_runtime_compiles_and_links(tmp_path):
    """The C++ runtime headers compile warning-free and a program using them
    links against the library (it is run by the GPU tests)."""
    exe = tmp_path / "runtime_sample"
    p = subprocess.run(["g++", "-std=c++17", "-O0", "-Wall", "-Wextra", "-Werror",
                        "-I", str(ROOT / "include"), str(ROOT / "tests" / "cpp" / "runtime_sample.cpp"),
                        "-o", str(exe), "-L", str(ROOT / "hipacc_amd"), "-lhipacc_gfx950"],
                       capture_output=True, text=True)
    assert p.returncode == 0, p.stderr


def test_adjoint_ring_and_band_enumerations(tmp_path):
    """The halo ring and border band the adjoint kernels enumerate (host
    build of the same header functions)."""
    exe = tmp_path / "ring_check"
    p = subprocess.run(["/opt/rocm/bin/hipcc", "--offload-arch=gfx950", "-O1", "-std=c++17",
                        "-I", str(ROOT / "include"), str(ROOT / "tests" / "cpp" / "ring_check.hip"),
                        "-o", str(exe)], capture_output=True, text=True)
    assert p.returncode == 0, p.stderr[-3000:]
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
