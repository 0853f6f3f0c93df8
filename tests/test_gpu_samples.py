"""The reference's own sample programs (samples-public/*/src/main.cpp),
unchanged, rewritten by the emitter and built against the runtime
(tests/samples/Makefile, from the reference checkout), run on the GPU.

Each sample that checks its result against its own CPU implementation
(compare_results in samples-public/common/hipacc_helper.hpp: integer pixels
within 1, float pixels within 0.1 %, over the image without the mask's
border) must print "Test PASSED"; the others must run to completion.  Their
input is the helper's synthetic data (no image library here).
"""
import subprocess
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]
BIN = ROOT / "tests" / "samples" / "bin"

# The samples that compare their result with their own CPU implementation.
SELF_CHECKING = {
    "Bilateral_Filter", "Bilateral_Filter_RGBA", "Box_Blur", "Box_Blur_RGBA",
    "Color_Conversion", "Dilate", "Dilate_RGBA", "Erode", "Erode_RGBA", "Gaussian_Blur",
    "Gaussian_Blur_RGBA", "Gaussian_Laplacian_Pyramid", "Histogram", "Kernel_Fusion_L2L",
    "Kernel_Fusion_L2P", "Kernel_Fusion_Mixed", "Kernel_Fusion_P2L", "Kernel_Fusion_P2P",
    "Laplace", "Laplace_RGBA", "Reduction_Max", "Reduction_Sum", "Scaling", "Sobel",
    "Sobel_RGBA", "Windowing",
}
# The others run to completion.
RUN_ONLY = {
    "Bokeh_Effect", "Chromatic_Abberation", "Game_of_Life", "Harris_Corner", "Mandelbrot",
    "Motion_Interpolation", "Night_Filter", "Optical_Flow", "ShiTomasi_Corner", "Unsharp",
    "WCE_Enhance",
}


@pytest.mark.parametrize("name", sorted(SELF_CHECKING | RUN_ONLY))
def test_reference_sample_runs(gpu, name, tmp_path):
    exe = BIN / name
    if not exe.exists():
        pytest.skip(f"{exe} is not built (make -C tests/samples, from the reference checkout)")
    print(f"sample {name}", flush=True)
    r = subprocess.run([str(exe)], cwd=tmp_path, capture_output=True, text=True, timeout=240)
    out = r.stdout + r.stderr
    print(out[-2000:])
    assert r.returncode == 0, f"{name} exited with {r.returncode}"
    assert "ERROR" not in out, f"{name} reported a runtime error"
    if name in SELF_CHECKING:
        assert "Test PASSED" in out and "Test FAILED" not in out
