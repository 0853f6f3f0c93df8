"""Mask rows and columns for tests: seeded, asymmetric, and different for the
row and the column pass, so that a transposed or mirrored mask, or a row and
column pass swapped, changes the result."""
from __future__ import annotations

import numpy as np


def skewed_1d(k: int, seed: int, signed: bool = False) -> np.ndarray:
    """k float32 taps.  signed=False: positive, normalised to sum 1 (keeps a
    rounded uint8 output in range); signed=True: standard normal."""
    rng = np.random.default_rng(seed)
    if signed:
        return rng.standard_normal(k).astype(np.float32)
    w = rng.uniform(0.2, 1.0, k)
    return (w / w.sum()).astype(np.float32)
