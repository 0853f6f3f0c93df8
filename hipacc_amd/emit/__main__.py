"""python -m hipacc_amd.emit prog.cpp -o out.cpp -I <dsl dir> [-D NAME[=V]] [--no-fp-contract]"""
import argparse
import sys

from . import EmitError, emit


def main() -> int:
    ap = argparse.ArgumentParser(prog="python -m hipacc_amd.emit",
                                 description="Rewrite a Hipacc DSL program for the MI355X runtime.")
    ap.add_argument("source")
    ap.add_argument("-o", "--output", required=True)
    ap.add_argument("-I", dest="includes", action="append", default=[],
                    help="include directory (the DSL headers' directory, hipacc.hpp)")
    ap.add_argument("-D", dest="defines", action="append", default=[])
    ap.add_argument("--no-fp-contract", action="store_true",
                    help="evaluate weighted sums as multiply then add (the DSL host emulation's "
                         "order) instead of fused multiply-adds (nvcc's default -fmad=true)")
    a = ap.parse_args()
    try:
        text = emit(a.source, a.includes, a.defines, fp_contract=not a.no_fp_contract)
    except EmitError as e:
        print(f"{a.source}: {e}", file=sys.stderr)
        return 1
    with open(a.output, "w") as f:
        f.write(text)
    return 0


if __name__ == "__main__":
    sys.exit(main())
