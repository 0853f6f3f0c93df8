"""Writes tests/golden/emit_golden.json: the outputs of the DSL programs
tests/dsl/*.cpp on the reference's own DSL host emulation
(oracle/_ref/dsl_<name>, built by `make -C oracle ref` from the reference
checkout's dsl/ headers), as SHA-256 digests of the output files and, for
the reductions, the values.  The emitted versions (tests/dsl/emitted/) must
reproduce them on the GPU (tests/test_gpu_emit.py).  Run from the repository
root:

    make -C oracle ref && python tests/golden/make_emit_golden.py
"""
from __future__ import annotations

import hashlib
import json
import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
PROGS = sorted(p.stem for p in (ROOT / "tests" / "dsl").glob("*.cpp"))


def main() -> int:
    out = {}
    with tempfile.TemporaryDirectory() as td:
        for name in PROGS:
            exe = ROOT / "oracle" / "_ref" / f"dsl_{name}"
            if not exe.exists():
                print(f"{exe} missing: run `make -C oracle ref` first", file=sys.stderr)
                return 1
            path = Path(td) / f"{name}.bin"
            subprocess.run([str(exe), str(path)], check=True)
            data = path.read_bytes()
            entry = {"bytes": len(data), "sha256": hashlib.sha256(data).hexdigest()}
            if name.startswith("reduce"):
                entry["values"] = [float(v) for v in np.frombuffer(data[:12], np.float32)]
                entry["f64"] = [float(v) for v in np.frombuffer(data[12:], np.float64)]
            out[name] = entry
    (HERE / "emit_golden.json").write_text(json.dumps(out, indent=1) + "\n")
    print(f"wrote {len(out)} programs")
    return 0


if __name__ == "__main__":
    sys.exit(main())
