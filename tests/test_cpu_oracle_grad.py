"""Pins the oracle's adjoints (oracle_local_grad, oracle_bilateral_grad) to
the forward they differentiate: FP64 central finite differences of the FP64
forward oracles (oracle_local_f64, oracle_bilateral_f64) on small images,
under every boundary mode, at every pixel (borders included, where the
boundary folding lives) and for every mask weight / the range sigma.

The convolution is linear in X and w, so its differences are exact up to
double rounding (tolerance 1e-9 relative); the bilateral filter is smooth and
the central difference has an O(h^2) error (h = 2^-10 on X, 2^-12 on sigma_r:
tolerance 2e-5 of the largest gradient)."""
import numpy as np
import pytest

import oracle_binding as ob
from hipacc_amd import capi

MODES = [capi.UNDEFINED, capi.CLAMP, capi.REPEAT, capi.MIRROR, capi.CONSTANT]


def conv_loss(x, w, g, kw, kh, mode, cval):
    out = np.zeros(x.shape, np.float64)
    ob.local(x, np.zeros_like(x), kw, kh, weights=w, boundary=mode, const_val=cval, f64_out=out)
    return float(np.sum(out * g))


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("kw,kh", [(3, 3), (5, 3), (5, 5)])
def test_conv_grad_matches_finite_differences(mode, kw, kh):
    rng = np.random.default_rng(10 * mode + kw + kh)
    H, W = 7, 9
    x = rng.random((H, W), dtype=np.float32)
    g = rng.standard_normal((H, W)).astype(np.float32)
    w = rng.standard_normal(kw * kh).astype(np.float32)
    cval = 0.75 if mode == capi.CONSTANT else 0.0
    dx, dw = ob.local_grad(x, g, w, kw, kh, boundary=mode, const_val=cval)
    h = np.float32(2.0 ** -10)
    fd = np.zeros_like(dx)
    for iy in range(H):
        for ix in range(W):
            xp, xm = x.copy(), x.copy()
            xp[iy, ix] += h
            xm[iy, ix] -= h
            step = float(xp[iy, ix]) - float(xm[iy, ix])
            fd[iy, ix] = (conv_loss(xp, w, g, kw, kh, mode, cval) -
                          conv_loss(xm, w, g, kw, kh, mode, cval)) / step
    assert np.allclose(dx, fd, rtol=1e-9, atol=1e-9 * np.abs(fd).max())
    fdw = np.zeros_like(dw)
    for t in range(kw * kh):
        wp, wm = w.copy(), w.copy()
        wp[t] += h
        wm[t] -= h
        step = float(wp[t]) - float(wm[t])
        fdw[t] = (conv_loss(x, wp, g, kw, kh, mode, cval) -
                  conv_loss(x, wm, g, kw, kh, mode, cval)) / step
    assert np.allclose(dw, fdw, rtol=1e-9, atol=1e-9 * np.abs(fdw).max())


def test_conv_grad_interior_is_flipped_correlation():
    """Away from the borders dX is the correlation of dY with the flipped
    mask (a closed form independent of the scatter loop)."""
    rng = np.random.default_rng(5)
    H, W, k = 12, 14, 5
    x = rng.random((H, W), dtype=np.float32)
    g = rng.standard_normal((H, W)).astype(np.float32)
    w = rng.standard_normal(k * k).astype(np.float32)
    dx, _ = ob.local_grad(x, g, w, k, k, boundary=capi.MIRROR)
    wf = w.reshape(k, k)[::-1, ::-1].astype(np.float64)
    r = k // 2
    for y in range(2 * r, H - 2 * r):
        for xx in range(2 * r, W - 2 * r):
            ref = np.sum(wf * g[y - r:y + r + 1, xx - r:xx + r + 1].astype(np.float64))
            assert abs(dx[y, xx] - ref) <= 1e-12 * max(1.0, abs(ref))


def bilat_loss(x, g, k, sd, sr, mode, cval):
    out = np.zeros(x.shape, np.float64)
    ob.bilateral(x, np.zeros_like(x), k, sd, sr, boundary=mode, const_val=cval, f64_out=out)
    return float(np.sum(out * g))


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("k", [3, 5])
def test_bilateral_grad_matches_finite_differences(mode, k):
    rng = np.random.default_rng(100 + 10 * mode + k)
    H, W = 8, 7
    x = rng.random((H, W), dtype=np.float32)
    g = rng.standard_normal((H, W)).astype(np.float32)
    sd, sr = 1.5, 0.3
    cval = 0.4 if mode == capi.CONSTANT else 0.0
    dx, ds = ob.bilateral_grad(x, g, k, sd, sr, boundary=mode, const_val=cval)
    h = np.float32(2.0 ** -10)
    fd = np.zeros_like(dx)
    for iy in range(H):
        for ix in range(W):
            xp, xm = x.copy(), x.copy()
            xp[iy, ix] += h
            xm[iy, ix] -= h
            step = float(xp[iy, ix]) - float(xm[iy, ix])
            fd[iy, ix] = (bilat_loss(xp, g, k, sd, sr, mode, cval) -
                          bilat_loss(xm, g, k, sd, sr, mode, cval)) / step
    assert np.max(np.abs(dx - fd)) <= 2e-5 * np.abs(fd).max()
    hs = 2.0 ** -12
    srp, srm = np.float32(sr + hs), np.float32(sr - hs)
    fds = (bilat_loss(x, g, k, sd, float(srp), mode, cval) -
           bilat_loss(x, g, k, sd, float(srm), mode, cval)) / (float(srp) - float(srm))
    assert abs(ds - fds) <= 2e-5 * max(abs(fds), 1e-3)
