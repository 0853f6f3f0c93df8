"""The C restatement (oracle/hipacc_oracle.c) against the reference's own DSL
host emulation (golden vectors from tests/golden/make_golden.py).

Every case is bit-exact: same tap order, same seeding, no FMA contraction.
"""
import numpy as np
import pytest

from hipacc_amd import capi

import golden_cases as gc
import oracle_binding as ob

CASES = gc.load()


@pytest.mark.parametrize("case,expected", CASES, ids=[gc.case_id(c) for c, _ in CASES])
def test_oracle_matches_reference_dsl(case, expected):
    op = case["op"]
    src = gc.input_for(case)
    spec = gc.local_spec(case)
    if spec is not None:
        out_dtype = spec.pop("out_dtype")
        dst = np.zeros(expected.shape, dtype=out_dtype)
        ob.local(src, dst, **spec)
        np.testing.assert_array_equal(dst, expected)
    elif op == "bilat_f32":
        k, sd, sr = int(case["args"][0]), float(case["args"][1]), float(case["args"][2])
        dst = np.zeros_like(src)
        ob.bilateral(src, dst, k, sd, sr, boundary=case["boundary"])
        np.testing.assert_array_equal(dst, expected)
    elif op == "red_f32":
        mode = {0: capi.SUM, 1: capi.MIN, 2: capi.MAX}[int(case["args"][0])]
        r = ob.reduce(src, mode, nthreads=1)
        assert np.float32(r) == expected[0]
    elif op == "red_u8":
        mode = {1: capi.MIN, 2: capi.MAX}[int(case["args"][0])]
        r = ob.reduce(src, mode, acc_dtype=capi.U32, nthreads=1)
        assert r == expected[0]
    elif op == "hist_u8":
        bins = ob.histogram(src, int(case["args"][0]))
        np.testing.assert_array_equal(bins, expected)
    elif op == "sub_i8":
        a, b = src
        dst = np.zeros_like(a)
        ob.point(capi.POINT_SUB, a, dst, b)
        np.testing.assert_array_equal(dst, expected)
    elif op in ("nn_f32", "lf_f32"):
        dw, dh = int(case["args"][0]), int(case["args"][1])
        dst = np.zeros((dh, dw), np.float32)
        interp = capi.INTERP_NN if op == "nn_f32" else capi.INTERP_LF
        ob.resample(src, dst, interp, boundary=case["boundary"])
        np.testing.assert_array_equal(dst, expected)
    elif op in ("interp_f32", "interp_u8"):
        interp, shape, kw = gc.interp_spec(case)
        dst = np.zeros(shape, src.dtype)
        ob.resample(src, dst, interp, **kw)
        np.testing.assert_array_equal(dst, expected)
    elif op in ("fuse_u8", "fuse_f32"):
        np.testing.assert_array_equal(ob.chain(src, gc.fuse_spec(case)), expected)
    else:
        pytest.fail(f"unmapped case {op}")
