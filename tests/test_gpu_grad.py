"""GPU adjoints against the oracle: the learnable-convolution backward
(hipacc_hip_local_grad) and the bilateral backward with a learnable range
sigma (hipacc_hip_bilateral_train / _grad), under every boundary mode, on
batched odd-sized images, plus the libtorch autograd Functions built on them.

Tolerance (float32 results against the double oracle): |gpu - ref| <= 1e-5 *
S, with S the sum of the absolute values of the terms that make up the
result (returned by the oracle); this is the float32 bound for a sum whose
terms cancel.  At full size (16 x 1024 x 1024) the adjoint identities
<conv(X), G> = <X, dX> = SUM_t w[t] dw[t] are checked instead."""
import numpy as np
import pytest
import torch

import oracle_binding as ob
from gpu_util import to_gpu, to_host
from hipacc_amd import capi, operators as op

pytestmark = pytest.mark.gpu

MODES = [capi.UNDEFINED, capi.CLAMP, capi.REPEAT, capi.MIRROR, capi.CONSTANT]
TOL = 1e-5


def _within(got, ref, scale, what):
    err = np.abs(got.astype(np.float64) - ref)
    bound = TOL * (scale + 1e-30)
    bad = err > bound
    assert not bad.any(), (f"{what}: {bad.sum()} of {bad.size} beyond tolerance; worst "
                           f"err {err.max():.3e} (bound there {bound.flat[np.argmax(err)]:.3e})")


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("kw,kh", [(3, 3), (5, 5), (7, 7), (9, 9), (5, 3), (11, 11)])
@pytest.mark.parametrize("shape", [(2, 37, 29), (1, 70, 131)])
def test_conv_grad_matches_oracle(gpu, mode, kw, kh, shape):
    print(f"conv_grad mode={mode} k={kw}x{kh} shape={shape}", flush=True)
    rng = np.random.default_rng(kw * 100 + kh * 10 + mode)
    x = ob.random_image(shape, np.float32, 11 + mode)
    g = rng.standard_normal(shape).astype(np.float32)
    w = (0.1 * rng.standard_normal(kw * kh)).astype(np.float32)
    cval = 0.5 if mode == capi.CONSTANT else 0.0
    dx, dw, dxa, dwa = ob.local_grad(x, g, w, kw, kh, boundary=mode, const_val=cval, with_abs=True)
    X, G = to_gpu(x), to_gpu(g)
    GX = torch.empty_like(X)
    W = torch.from_numpy(w).cuda()
    GW = torch.empty(kw * kh, dtype=torch.float32, device="cuda")
    op.prepare_local_grad(X, G, GX, kw, kh, weights_dev=W, grad_weights=GW, boundary=mode,
                          const_val=cval).launch()
    _within(to_host(GX, np.float32), dx, dxa, "dX")
    _within(to_host(GW, np.float32), dw, dwa, "dw")
    # host weights, each output alone
    GX2 = torch.empty_like(X)
    op.prepare_local_grad(X, G, GX2, kw, kh, weights=w, boundary=mode, const_val=cval).launch()
    assert torch.equal(GX, GX2)
    GW2 = torch.empty_like(GW)
    op.prepare_local_grad(X, G, None, kw, kh, weights=w, grad_weights=GW2, boundary=mode,
                          const_val=cval).launch()
    assert torch.equal(GW, GW2), "dw differs between runs (must be deterministic)"
    for variant in (capi.VAR_TILE, capi.VAR_STREAM):
        GXv, GWv = torch.empty_like(X), torch.empty_like(GW)
        op.prepare_local_grad(X, G, GXv, kw, kh, weights_dev=W, grad_weights=GWv, boundary=mode,
                              const_val=cval).with_variant(variant).launch()
        assert torch.equal(GXv, GX)
        _within(to_host(GWv, np.float32), dw, dwa, f"dw variant {variant}")


@pytest.mark.parametrize("mode", [capi.CLAMP, capi.MIRROR, capi.REPEAT])
def test_conv_grad_tiny_images(gpu, mode):
    """Images narrower than the halo: the folding reaches across the image."""
    for shape, k in (((1, 3, 2), 9), ((1, 1, 5), 5), ((2, 4, 4), 7)):
        rng = np.random.default_rng(7)
        x = rng.random(shape, dtype=np.float32)
        g = rng.standard_normal(shape).astype(np.float32)
        w = rng.standard_normal(k * k).astype(np.float32)
        dx, dw, dxa, dwa = ob.local_grad(x, g, w, k, k, boundary=mode, with_abs=True)
        X, G = to_gpu(x), to_gpu(g)
        GX = torch.empty_like(X)
        GW = torch.empty(k * k, dtype=torch.float32, device="cuda")
        op.prepare_local_grad(X, G, GX, k, k, weights=w, grad_weights=GW, boundary=mode).launch()
        _within(to_host(GX, np.float32), dx, dxa, f"dX {shape} k{k}")
        _within(to_host(GW, np.float32), dw, dwa, f"dw {shape} k{k}")


@pytest.mark.parametrize("mode", [capi.CLAMP, capi.MIRROR])
def test_conv_adjoint_identity_full_size(gpu, mode):
    """16 x 1024 x 1024, 9x9 learnable mask: <Y, G> = <X, dX> = <w, dw>."""
    B, H, W, k = 16, 1024, 1024, 9
    X = torch.empty((B, H, W), dtype=torch.float32, device="cuda")
    G = torch.empty_like(X)
    op.fill_random(X, 41)
    op.fill_random(G, 42)
    G -= 0.5
    w = (0.1 * np.random.default_rng(3).standard_normal(k * k)).astype(np.float32)
    Wd = torch.from_numpy(w).cuda()
    Y = torch.empty_like(X)
    op.local(X, Y, k, k, weights_dev=Wd, boundary=mode)
    GX = torch.empty_like(X)
    GW = torch.empty(k * k, dtype=torch.float32, device="cuda")
    op.prepare_local_grad(X, G, GX, k, k, weights_dev=Wd, grad_weights=GW, boundary=mode).launch()
    torch.cuda.synchronize()
    yg = torch.sum(Y.double() * G.double()).item()
    xdx = torch.sum(X.double() * GX.double()).item()
    wdw = float(np.dot(w.astype(np.float64), GW.cpu().double().numpy()))
    scale = torch.sum(torch.abs(Y.double() * G.double())).item()
    assert abs(yg - xdx) <= 1e-5 * scale, (yg, xdx, scale)
    assert abs(yg - wdw) <= 1e-5 * scale, (yg, wdw, scale)


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("k", [3, 5, 7, 9, 15])
@pytest.mark.parametrize("shape", [(2, 37, 29), (1, 70, 131)])
def test_bilateral_grad_matches_oracle(gpu, mode, k, shape):
    print(f"bilateral_grad mode={mode} k={k} shape={shape}", flush=True)
    rng = np.random.default_rng(k * 10 + mode)
    x = ob.random_image(shape, np.float32, 21 + mode)
    g = rng.standard_normal(shape).astype(np.float32)
    sd, sr = 2.0, 0.1
    cval = 0.3 if mode == capi.CONSTANT else 0.0
    dx, ds, dxa, dsa = ob.bilateral_grad(x, g, k, sd, sr, boundary=mode, const_val=cval,
                                         with_abs=True)
    X, G = to_gpu(x), to_gpu(g)
    Y, D, GX = torch.empty_like(X), torch.empty_like(X), torch.empty_like(X)
    GS = torch.zeros(1, dtype=torch.float32, device="cuda")
    op.prepare_bilateral_train(X, Y, D, k, sd, sr, boundary=mode, const_val=cval).launch()
    op.prepare_bilateral_grad(X, Y, D, G, GX, k, sd, sr, grad_sigma=GS, boundary=mode,
                              const_val=cval).launch()
    _within(to_host(GX, np.float32), dx, dxa, "dX")
    _within(to_host(GS, np.float32), np.array([ds]), np.array([dsa]), "dsigma_r")
    # the training forward equals the inference forward
    Y2 = torch.empty_like(X)
    op.bilateral(X, Y2, k, sd, sr, boundary=mode, const_val=cval)
    assert torch.equal(Y, Y2)


def test_autograd_functions(gpu):
    """The libtorch autograd Functions return the C-ABI gradients."""
    from hipacc_amd import autograd as ag
    rng = np.random.default_rng(9)
    x = ob.random_image((2, 40, 52), np.float32, 5)
    g = rng.standard_normal(x.shape).astype(np.float32)
    w = (0.1 * rng.standard_normal(81)).astype(np.float32)
    X = torch.from_numpy(x).cuda().requires_grad_(True)
    Wt = torch.from_numpy(w.reshape(9, 9)).cuda().requires_grad_(True)
    Y = ag.conv2d(X, Wt, capi.MIRROR)
    Y.backward(torch.from_numpy(g).cuda())
    dx, dw, dxa, dwa = ob.local_grad(x, g, w, 9, 9, boundary=capi.MIRROR, with_abs=True)
    _within(X.grad.detach().cpu().numpy(), dx, dxa, "autograd dX")
    _within(Wt.grad.detach().cpu().numpy().ravel(), dw, dwa, "autograd dw")

    X2 = torch.from_numpy(x).cuda().requires_grad_(True)
    S = torch.tensor(0.1, device="cuda", requires_grad=True)
    Y2 = ag.bilateral(X2, S, 7, 2.0, capi.CLAMP)
    Y2.backward(torch.from_numpy(g).cuda())
    bdx, bds, bdxa, bdsa = ob.bilateral_grad(x, g, 7, 2.0, 0.1, boundary=capi.CLAMP,
                                             with_abs=True)
    _within(X2.grad.detach().cpu().numpy(), bdx, bdxa, "autograd bilateral dX")
    _within(np.array([S.grad.item()]), np.array([bds]), np.array([bdsa]), "autograd dsigma")
