"""The batched pyramid on the geometry where the plan runs levels on wide
(128 x 32) tiles: at least 512 of them at a level.  Bit-exact against the
oracle's per-level kernel sequence, with and without contraction, CLAMP and
MIRROR: 8 levels over large and odd-sized images, and a 4-level set where
images of 8 and 9 rows share a wide level with 2048^2 images."""
import numpy as np
import pytest
import torch

import oracle_binding as ob
from gpu_util import to_gpu, to_host
from hipacc_amd import capi, operators as op

# (H, W) at level 0
LARGE = [(2048, 2048), (2161, 4097), (1151, 1160), (1999, 1001)]
LARGE_LEVELS = 8
MIXED = [(2048, 2048), (2048, 2048), (17, 600), (19, 37), (18, 18), (33, 9)]
MIXED_LEVELS = 4


def _run(shapes, levels, mode, fma, seed):
    gauss, lap, host = [], [], []
    for i, (h, w) in enumerate(shapes):
        g0 = ob.random_image((h, w), np.float32, seed + i)
        host.append(g0)
        gauss.append([to_gpu(g0)] + [torch.full((h >> l, w >> l), -1.0, dtype=torch.float32,
                                                device="cuda") for l in range(1, levels)])
        lap.append([torch.full((h >> l, w >> l), -1.0, dtype=torch.float32, device="cuda")
                    for l in range(levels - 1)])
    plan = op.PyramidPlan(gauss, lap, boundary=mode, fp_contract=fma)
    plan.run()
    torch.cuda.synchronize()
    for i, g0 in enumerate(host):
        G, L = ob.pyramid(g0, levels, boundary=mode, fp_contract=fma)
        for l in range(1, levels):
            np.testing.assert_array_equal(to_host(gauss[i][l], np.float32), G[l],
                                          err_msg=f"image {i} {shapes[i]} G[{l}]")
        for l in range(levels - 1):
            np.testing.assert_array_equal(to_host(lap[i][l], np.float32), L[l],
                                          err_msg=f"image {i} {shapes[i]} L[{l}]")
    plan.close()


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [capi.CLAMP, capi.MIRROR])
@pytest.mark.parametrize("fma", [True, False])
def test_wide_tiles_8_levels(gpu, mode, fma):
    narrow, wide = op.pyramid_plan_tiles(LARGE, LARGE_LEVELS)
    assert wide[0] >= 512, wide   # level 0 runs on the wide-tile kernel
    print(f"wide-tile pyramid 8 levels mode={mode} fma={fma} wide={wide}", flush=True)
    _run(LARGE, LARGE_LEVELS, mode, fma, 900)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [capi.CLAMP, capi.MIRROR])
@pytest.mark.parametrize("fma", [True, False])
def test_wide_tiles_mixed_threshold(gpu, mode, fma):
    narrow, wide = op.pyramid_plan_tiles(MIXED, MIXED_LEVELS)
    assert wide[1] >= 512 and narrow[1] > 0, (narrow, wide)
    print(f"wide-tile pyramid mixed mode={mode} fma={fma} wide={wide} narrow={narrow}", flush=True)
    _run(MIXED, MIXED_LEVELS, mode, fma, 950)
