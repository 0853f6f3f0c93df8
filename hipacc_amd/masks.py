"""Mask tables used by the workloads, computed the same way everywhere
(tests, benchmark, reference-DSL driver): in double precision, normalised,
then rounded once to float32."""
from __future__ import annotations

import math

import numpy as np


def gaussian_1d(k: int, sigma: float) -> np.ndarray:
    """Normalised 1D Gaussian, w(i) = exp(-i^2 / (2 sigma^2)) / sum."""
    r = k // 2
    t = [math.exp(-(i * i) / (2.0 * sigma * sigma)) for i in range(-r, r + 1)]
    s = 0.0
    for v in t:  # sequential double sum, as the C++ driver's loop
        s += v
    return np.array([v / s for v in t]).astype(np.float32)


def gaussian_2d(k: int, sigma: float) -> np.ndarray:
    """Normalised 2D Gaussian (k x k), as the reference-DSL driver builds it:
    w(i,j) = exp(-(i^2+j^2)/(2 sigma^2)) / sum over the window."""
    r = k // 2
    t = [[math.exp(-(x * x + y * y) / (2.0 * sigma * sigma)) for x in range(-r, r + 1)]
         for y in range(-r, r + 1)]
    s = 0.0
    for row in t:  # sequential row-major double sum, as the C++ driver's loop
        for v in row:
            s += v
    return np.array([[v / s for v in row] for row in t]).astype(np.float32)


def bilateral_spatial(k: int, sigma_d: float) -> np.ndarray:
    """Spatial weights of the bilateral filter, exp(-(dx^2+dy^2)/(2 sigma_d^2)),
    unnormalised (the filter divides by the weight sum)."""
    r = k // 2
    t = np.array([[math.exp(-(x * x + y * y) / (2.0 * sigma_d * sigma_d)) for x in range(-r, r + 1)]
                  for y in range(-r, r + 1)])
    return t.astype(np.float32)


SOBEL_X = np.array([[-1, 0, 1], [-2, 0, 2], [-1, 0, 1]], dtype=np.int32)
SOBEL_Y = np.array([[-1, -2, -1], [0, 0, 0], [1, 2, 1]], dtype=np.int32)
BINOMIAL5 = np.array([1, 4, 6, 4, 1], dtype=np.float32) / np.float32(16.0)
