"""The emitted host programs (tests/dsl/emitted/*.cpp, built into
tests/dsl/bin/ against the C++ runtime) on the GPU against the same DSL
programs on the reference's DSL host emulation: the digests in
tests/golden/emit_golden.json (and, where oracle/_ref/dsl_<name> is built,
that program's output, for the report of a mismatch).

Bar: bit-exact for the images (the programs are emitted and compiled
without contraction, the DSL host emulation's arithmetic), except programs
calling transcendental functions in their own kernels (device expf against
glibc's: within TOLERANCE of the DSL program's output); the float sum of the
reduction within 1e-6 of the sum of magnitudes of the double sum (the
emulation's sequential float sum differs from both by up to its own
n * eps bound), its minimum and maximum exact.
"""
import hashlib
import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]
GOLDEN = json.loads((ROOT / "tests" / "golden" / "emit_golden.json").read_text())
# Programs whose kernels call transcendental functions (device expf against
# glibc's): pixel type and the largest difference allowed per pixel.
TOLERANCE = {"bilateral_u8": (np.uint8, 1)}


def _run(exe: Path, out: Path) -> bytes:
    print(f"run {exe.name}", flush=True)
    subprocess.run([str(exe), str(out)], check=True, timeout=120)
    return out.read_bytes()


@pytest.mark.parametrize("name", sorted(GOLDEN))
def test_emitted_program_matches_dsl(gpu, name, tmp_path):
    exe = ROOT / "tests" / "dsl" / "bin" / name
    assert exe.exists(), f"{exe} is not built (make -C tests/dsl)"
    data = _run(exe, tmp_path / "gpu.bin")
    g = GOLDEN[name]
    assert len(data) == g["bytes"]
    if "values" in g:
        got = np.frombuffer(data[:12], np.float32)
        exp = np.array(g["values"], np.float32)
        exact, magnitude, n = np.frombuffer(data[12:], np.float64)
        assert got[0] == exp[0] and got[1] == exp[1]
        # the GPU's tree sum against the double sum, 1e-6 of the sum of
        # magnitudes; the emulation's sequential float sum only within its
        # own n * eps bound
        assert abs(float(got[2]) - exact) <= 1e-6 * magnitude
        assert abs(float(exp[2]) - exact) <= n * 2.0 ** -24 * magnitude
        return
    if hashlib.sha256(data).hexdigest() == g["sha256"]:
        return
    ref_exe = ROOT / "oracle" / "_ref" / f"dsl_{name}"
    detail = ""
    if ref_exe.exists():
        ref = _run(ref_exe, tmp_path / "dsl.bin")
        assert hashlib.sha256(ref).hexdigest() == g["sha256"], "the DSL program changed"
        if name in TOLERANCE:
            dt, tol = TOLERANCE[name]
            a = np.frombuffer(data, dt).astype(np.int64)
            b = np.frombuffer(ref, dt).astype(np.int64)
            diff = np.abs(a - b)
            print(f"{name}: {np.count_nonzero(diff)} pixels differ, max {diff.max()}")
            assert diff.max() <= tol
            return
        a, b = np.frombuffer(data, np.uint8), np.frombuffer(ref, np.uint8)
        bad = np.nonzero(a != b)[0]
        detail = f": {bad.size} bytes differ, first at byte {bad[0] if bad.size else -1}"
    pytest.fail(f"{name} output differs from the DSL emulation{detail}")
