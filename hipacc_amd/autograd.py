"""Differentiable operators: libtorch autograd Functions (compiled from
torch_ext/hipacc_autograd.cpp) whose forward and backward are the gfx950
kernels of libhipacc_gfx950.so.

    y = conv2d(x, w, boundary)                  learnable (kh, kw) float mask
    y = bilateral(x, sigma_r, ksize, sigma_d, boundary)   learnable range sigma

x is a float32 (H, W) or (B, H, W) device tensor; boundary is one of
capi.UNDEFINED / CLAMP / REPEAT / MIRROR / CONSTANT.  The extension is built
in-tree by hipacc_amd.build.build_torch_ext(); importing this module without
it raises ImportError (there is no fallback path).
"""
from __future__ import annotations

import importlib.machinery
import importlib.util

import torch  # noqa: F401  (loads the libraries the extension links against)

from . import build as _build
from . import capi

_path = _build.torch_ext_path()
if not _path.exists():
    raise ImportError(f"{_path} is not built; run hipacc_amd.build.build_torch_ext()")
capi.lib()  # the kernel library, from the package directory
_loader = importlib.machinery.ExtensionFileLoader("hipacc_autograd", str(_path))
_spec = importlib.util.spec_from_file_location("hipacc_autograd", str(_path), loader=_loader)
_ext = importlib.util.module_from_spec(_spec)
_loader.exec_module(_ext)


def conv2d(x: torch.Tensor, w: torch.Tensor, boundary: int = capi.CLAMP,
           const_val: float = 0.0) -> torch.Tensor:
    """Y(p) = SUM_t w[t] * X_bh(p + t - r), differentiable in x and w."""
    return _ext.conv2d(x, w, int(boundary), float(const_val))


def bilateral(x: torch.Tensor, sigma_r: torch.Tensor, ksize: int, sigma_d: float,
              boundary: int = capi.MIRROR, const_val: float = 0.0) -> torch.Tensor:
    """Bilateral filter differentiable in x and the range sigma (a one-element
    tensor; its value is read on the host at each forward)."""
    return _ext.bilateral(x, sigma_r, int(ksize), float(sigma_d), int(boundary), float(const_val))
