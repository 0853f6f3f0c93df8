"""Builds libhipacc_gfx950.so (the C-ABI kernel library) in-tree.

Every ``csrc/*.hip`` file is compiled for gfx950 with hipcc into
``hipacc_amd/_build/`` in parallel and linked into
``hipacc_amd/libhipacc_gfx950.so``.  Floating-point contraction is off
(``-ffp-contract=off``): kernels that want fused multiply-adds call
``__builtin_fmaf`` explicitly, so each kernel's rounding sequence is the one
its source states.  Packed FP32 instructions are disabled.

``build_torch_ext()`` compiles the libtorch autograd Functions
(``torch_ext/hipacc_autograd.cpp``) with g++ into
``hipacc_amd/hipacc_autograd*.so``, linked to the kernel library.

Run ``python -m hipacc_amd.build`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libhipacc_gfx950.so"
TORCH_EXT_SRC = PKG / "torch_ext" / "hipacc_autograd.cpp"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-I", str(PKG.parent / "include"),
]


def _compile(src: Path) -> Path:
    obj = BUILD / (src.stem + ".o")
    cmd = [HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src.name}:\n{res.stderr[-6000:]}")
    return obj


def build(jobs: int | None = None, verbose: bool = True) -> Path:
    """Compiles every kernel source and links the shared library."""
    BUILD.mkdir(exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip"))
    jobs = jobs or min(8, os.cpu_count() or 1)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stderr[-6000:]}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources")
    return LIB


def torch_ext_path() -> Path:
    import sysconfig
    return PKG / ("hipacc_autograd" + sysconfig.get_config_var("EXT_SUFFIX"))


def build_torch_ext(verbose: bool = True) -> Path:
    """Compiles the autograd extension module against the installed torch."""
    import sysconfig

    import torch
    from torch.utils import cpp_extension as ce

    out = torch_ext_path()
    tmp = out.with_suffix(".tmp")
    abi = int(torch._C._GLIBCXX_USE_CXX11_ABI)
    cmd = ["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-w",
           f"-D_GLIBCXX_USE_CXX11_ABI={abi}", "-DTORCH_EXTENSION_NAME=hipacc_autograd",
           "-DTORCH_API_INCLUDE_EXTENSION_H", "-D__HIP_PLATFORM_AMD__=1",
           *[f"-I{p}" for p in ce.include_paths()], "-I/opt/rocm/include",
           f"-I{sysconfig.get_paths()['include']}", f"-I{PKG.parent / 'include'}",
           str(TORCH_EXT_SRC), "-o", str(tmp),
           *[f"-L{p}" for p in ce.library_paths()], "-lc10", "-lc10_hip", "-ltorch", "-ltorch_cpu",
           "-ltorch_python", f"-L{PKG}", "-lhipacc_gfx950", "-Wl,-rpath,$ORIGIN"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"g++ failed on {TORCH_EXT_SRC.name}:\n{res.stderr[-6000:]}")
    os.replace(tmp, out)
    if verbose:
        print(f"built {out}")
    return out


if __name__ == "__main__":
    build(jobs=int(sys.argv[1]) if len(sys.argv) > 1 else None)
    build_torch_ext()
