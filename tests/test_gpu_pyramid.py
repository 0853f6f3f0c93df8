"""Resampling operators and batched pyramids on the GPU against the oracle,
bit-exact: hipacc_hip_resample (NN / LF, COPY / SUB / ADD, windows, boundary
modes), the batched pyramid plan (graph replay and direct launches) over
uneven images, and the plan against the per-image, per-level sequence of
library launches (convolution, NN subsample, LF difference)."""
import numpy as np
import pytest
import torch

import oracle_binding as ob
from gpu_util import to_gpu, to_host
from hipacc_amd import capi, operators as op

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("interp", [capi.INTERP_NN, capi.INTERP_LF])
@pytest.mark.parametrize("src_shape,dst_shape", [((29, 37), (14, 18)), ((14, 18), (29, 37)),
                                                 ((48, 64), (24, 32)), ((33, 130), (70, 61))])
@pytest.mark.parametrize("mode", [capi.CLAMP, capi.MIRROR, capi.REPEAT, capi.CONSTANT])
def test_resample_matches_oracle(gpu, interp, src_shape, dst_shape, mode):
    src = ob.random_image(src_shape, np.float32, 5)
    b = ob.random_image(dst_shape, np.float32, 6)
    for rop in (capi.RESAMPLE_COPY, capi.RESAMPLE_SUB, capi.RESAMPLE_ADD):
        ref = np.zeros(dst_shape, np.float32)
        ob.resample(src, ref, interp, op=rop, src2=b, boundary=mode, const_val=0.25)
        out = torch.empty(dst_shape, dtype=torch.float32, device="cuda")
        op.resample(to_gpu(src), out, interp, op=rop, src2=to_gpu(b), boundary=mode,
                    const_val=0.25)
        np.testing.assert_array_equal(to_host(out, np.float32), ref)


def test_resample_windows(gpu):
    src = ob.random_image((40, 50), np.float32, 8)
    ref = np.zeros((30, 30), np.float32)
    ob.resample(src, ref, capi.INTERP_LF, boundary=capi.MIRROR, src_window=(5, 3, 20, 17))
    out = torch.zeros((30, 30), dtype=torch.float32, device="cuda")
    op.resample(to_gpu(src), out, capi.INTERP_LF, boundary=capi.MIRROR, src_window=(5, 3, 20, 17))
    np.testing.assert_array_equal(to_host(out, np.float32), ref)


SHAPES = [(29, 37), (130, 67), (190, 257), (64, 64), (300, 129), (41, 1000)]


def _alloc(shapes, levels, seed):
    gauss, lap, host = [], [], []
    for i, (h, w) in enumerate(shapes):
        g0 = ob.random_image((h, w), np.float32, seed + i)
        host.append(g0)
        gl = [to_gpu(g0)]
        for l in range(1, levels):
            gl.append(torch.full((h >> l, w >> l), -1.0, dtype=torch.float32, device="cuda"))
        gauss.append(gl)
        lap.append([torch.full((h >> l, w >> l), -1.0, dtype=torch.float32, device="cuda")
                    for l in range(levels - 1)])
    return gauss, lap, host


def _check(gauss, lap, host, levels, **kw):
    for i, g0 in enumerate(host):
        G, L = ob.pyramid(g0, levels, **kw)
        for l in range(1, levels):
            np.testing.assert_array_equal(to_host(gauss[i][l], np.float32), G[l],
                                          err_msg=f"image {i} G[{l}]")
        for l in range(levels - 1):
            np.testing.assert_array_equal(to_host(lap[i][l], np.float32), L[l],
                                          err_msg=f"image {i} L[{l}]")


@pytest.mark.parametrize("tile_rows", [8, 16, 32])
@pytest.mark.parametrize("fma", [True, False])
def test_pyramid_plan_matches_oracle(gpu, tile_rows, fma):
    levels = 5
    gauss, lap, host = _alloc(SHAPES, levels, 100)
    print(f"pyramid plan tile_rows={tile_rows} fma={fma}", flush=True)
    plan = op.PyramidPlan(gauss, lap, tile_rows=tile_rows, fp_contract=fma)
    plan.run()
    torch.cuda.synchronize()
    _check(gauss, lap, host, levels, fp_contract=fma)
    plan.close()


@pytest.mark.parametrize("mode", [capi.MIRROR, capi.REPEAT, capi.CONSTANT])
def test_pyramid_plan_boundary_modes(gpu, mode):
    levels = 4
    gauss, lap, host = _alloc(SHAPES[:4], levels, 200)
    plan = op.PyramidPlan(gauss, lap, boundary=mode)
    plan.run_direct()
    torch.cuda.synchronize()
    _check(gauss, lap, host, levels, boundary=mode)
    plan.close()


def test_pyramid_graph_replays(gpu):
    """Each replay of the captured graph, with eager work and new inputs in
    between, gives the oracle's result."""
    levels = 6
    shapes = [(256, 300), (130, 67), (97, 500)]
    gauss, lap, host = _alloc(shapes, levels, 300)
    plan = op.PyramidPlan(gauss, lap)
    for rep in range(4):
        for i, (h, w) in enumerate(shapes):
            g0 = ob.random_image((h, w), np.float32, 400 + 10 * rep + i)
            host[i] = g0
            gauss[i][0].copy_(torch.from_numpy(g0))
        scratch = torch.randn(1 << 20, device="cuda")
        scratch.mul_(2.0)
        plan.run()
        torch.cuda.synchronize()
        _check(gauss, lap, host, levels)
    plan.close()


def test_pyramid_plan_equals_per_level_launches(gpu):
    """The batched plan and the sample's per-image, per-level kernel
    sequence run through the library give the same bits."""
    levels = 5
    gauss, lap, host = _alloc(SHAPES, levels, 500)
    plan = op.PyramidPlan(gauss, lap)
    plan.run()
    g5 = ob.BINOMIAL5
    for i, (h, w) in enumerate(SHAPES):
        G = [gauss[i][0]]
        for l in range(levels - 1):
            t = torch.empty_like(G[l])
            op.local(G[l], t, 5, 5, weights=g5, boundary=capi.CLAMP)
            g1 = torch.empty(((h >> (l + 1)), (w >> (l + 1))), dtype=torch.float32, device="cuda")
            op.resample(t, g1, capi.INTERP_NN)
            lp = torch.empty_like(G[l])
            op.resample(g1, lp, capi.INTERP_LF, op=capi.RESAMPLE_SUB, src2=G[l])
            assert torch.equal(g1, gauss[i][l + 1]), f"image {i} G[{l + 1}]"
            assert torch.equal(lp, lap[i][l]), f"image {i} L[{l}]"
            G.append(g1)
    plan.close()


# ---- the interpolation set: NN / LF / B5 / CF / L3 ------------------------------
import golden_cases as gc  # noqa: E402

INTERP_CASES = [(c, e) for c, e in gc.load() if c["op"] in ("interp_f32", "interp_u8")]


def _check_interp(got, want, interp, what):
    if interp == capi.INTERP_L3:
        # Lanczos weights are evaluated in double with the device's sin; the
        # weight rounded to float can differ from glibc's in its last bit
        if got.dtype == np.uint8:
            d = np.abs(got.astype(np.int32) - want.astype(np.int32))
            assert d.max() <= 1 and np.count_nonzero(d) <= max(2, got.size // 1000), what
        else:
            np.testing.assert_allclose(got, want, rtol=2e-6, atol=2e-6, err_msg=what)
    else:
        np.testing.assert_array_equal(got, want, err_msg=what)


@pytest.mark.parametrize("case,expected", INTERP_CASES,
                         ids=[gc.case_id(c) for c, _ in INTERP_CASES])
def test_interp_golden(gpu, case, expected):
    """GPU resampling against the reference DSL's own output."""
    src = gc.input_for(case)
    interp, shape, kw = gc.interp_spec(case)
    out = torch.zeros(shape, dtype=torch.uint8 if src.dtype == np.uint8 else torch.float32,
                      device="cuda")
    op.resample(to_gpu(src), out, interp, **kw)
    _check_interp(to_host(out, src.dtype), expected, interp, gc.case_id(case))


@pytest.mark.parametrize("interp", [capi.INTERP_NN, capi.INTERP_LF, capi.INTERP_B5,
                                    capi.INTERP_CF, capi.INTERP_L3])
@pytest.mark.parametrize("mode", [capi.CLAMP, capi.MIRROR, capi.REPEAT, capi.CONSTANT])
@pytest.mark.parametrize("dt", [np.float32, np.uint8])
def test_interp_vs_oracle(gpu, interp, mode, dt):
    """Down- and up-sampling of larger images, a cropped accessor window and
    call offsets, against the oracle."""
    print(f"interp {interp} mode {mode} {np.dtype(dt)}", flush=True)
    for shape_in, shape_out, win, xf, yf in (((300, 517), (121, 233), None, 0, 0),
                                             ((97, 130), (333, 517), None, 3, -2),
                                             ((200, 240), (150, 170), (17, 9, 180, 150), -1, 4)):
        src = ob.random_image(shape_in, dt, 70 + interp)
        ref = np.zeros(shape_out, dt)
        ob.resample(src, ref, interp, boundary=mode, const_val=7, src_window=win, xf=xf, yf=yf)
        out = torch.zeros(shape_out, dtype=torch.uint8 if dt == np.uint8 else torch.float32,
                          device="cuda")
        op.resample(to_gpu(src), out, interp, boundary=mode, const_val=7, src_window=win,
                    xf=xf, yf=yf)
        _check_interp(to_host(out, dt), ref, interp, f"{shape_in}->{shape_out} win {win}")
