"""Helpers for GPU tests: host <-> device copies of images, pitched buffers.

uint32 / uint16 images are held in int32 / int16 device tensors (same bytes);
the C ABI gets their pixel type through the explicit dtype argument.
"""
from __future__ import annotations

import numpy as np
import torch

_STORE = {np.dtype(np.uint8): (np.uint8, torch.uint8), np.dtype(np.int8): (np.int8, torch.int8),
          np.dtype(np.int16): (np.int16, torch.int16), np.dtype(np.uint16): (np.int16, torch.int16),
          np.dtype(np.int32): (np.int32, torch.int32), np.dtype(np.uint32): (np.int32, torch.int32),
          np.dtype(np.float32): (np.float32, torch.float32)}


def to_gpu(a: np.ndarray, pitch: int = 0) -> torch.Tensor:
    """Device copy of a (H, W) or (B, H, W) array; with pitch > W the rows are
    padded (the returned tensor is a column slice of a wider buffer)."""
    nps, tds = _STORE[a.dtype]
    t = torch.from_numpy(np.ascontiguousarray(a).view(nps)).cuda()
    if pitch and pitch > a.shape[-1]:
        shape = list(a.shape)
        shape[-1] = pitch
        buf = torch.zeros(shape, dtype=tds, device="cuda")
        buf[..., : a.shape[-1]] = t
        t = buf[..., : a.shape[-1]]
    return t


def empty_gpu(shape, np_dtype, pitch: int = 0, fill: int = 0) -> torch.Tensor:
    shape = list(shape)
    w = shape[-1]
    if pitch and pitch > w:
        shape[-1] = pitch
    _, tds = _STORE[np.dtype(np_dtype)]
    return torch.full(shape, fill, dtype=tds, device="cuda")[..., :w]


def to_host(t: torch.Tensor, np_dtype) -> np.ndarray:
    torch.cuda.synchronize()
    return t.contiguous().cpu().numpy().view(np.dtype(np_dtype))
