"""The HIP emitter (hipacc_amd.emit): operator recognition from kernel()
bodies, and the emitted host programs of tests/dsl/ reproduced exactly by a
fresh run of the emitter (where libclang and the DSL headers exist).
The GPU side (the emitted programs against the reference's DSL host
emulation) is tests/test_gpu_emit.py."""
import os
import re
from pathlib import Path

import pytest

from hipacc_amd.emit import recognise as R

ROOT = Path(__file__).resolve().parents[1]
DSL = ROOT / "tests" / "dsl"
DSL_HEADERS = "/root/reference/dsl"


def toks(s):
    return re.findall(r"\d+\.\d*f?|\.\d+f?|\d+f?|\w+|->|::|[-+*/]=|[-+*/=]|[(){}\[\],;&<>]", s)


def kernel(body, acc=("input",), masks=("mask",), doms=("dom",)):
    return R.classify_kernel(toks("{ " + body + " }"), list(acc), list(masks), list(doms))


def test_local_convolve_rounded():
    op = kernel("output() = (uchar)(convolve(mask, Reduce::SUM, [&] () -> float "
                "{ return mask() * input(mask); }) + 0.5f);")
    assert isinstance(op, R.LocalOp)
    assert (op.accessor, op.mask, op.domain, op.reduce, op.acc_type, op.round) == \
        ("input", "mask", None, "SUM", "float", True)


def test_local_convolve_operand_order_and_division():
    op = kernel("output() = convolve(mask, Reduce::SUM, [&] () -> int "
                "{ return input(mask) * mask(); }) / 16;")
    assert (op.mask, op.acc_type, op.post_div, op.round) == ("mask", "int", "16", False)


def test_local_reduce_domain_forms():
    op = kernel("output() = reduce(dom, Reduce::MIN, [&] () -> uchar { return input(dom); });")
    assert (op.mask, op.domain, op.reduce) == (None, "dom", "MIN")
    op = kernel("output() = (short)reduce(dom, Reduce::SUM, [&] () -> int "
                "{ return mask(dom) * input(dom); });")
    assert (op.mask, op.domain, op.acc_type) == ("mask", "dom", "int")


def test_point_forms():
    assert kernel("output() = input();").op == "COPY"
    op = kernel("output() = input() * 0.5f + 0.25f;")
    assert (op.op, op.alpha, op.beta) == ("AFFINE", "0.5f", "0.25f")
    op = kernel("output() = 2.0f * input() - 1;")
    assert (op.op, op.alpha, op.beta) == ("AFFINE", "2.0f", "-1")
    op = kernel("float v = input(); v += 3; output() = v;")
    assert (op.op, op.alpha, op.beta) == ("AFFINE", "1.0f", "3")
    op = kernel("output() = a() - b();", acc=("a", "b"))
    assert (op.op, op.inputs) == ("SUB", ["a", "b"])


@pytest.mark.parametrize("body", [
    "output() = (input() + 2) * 0.5f;",            # add then multiply: not alpha * x + beta
    "output() = input() / 3.0f;",                  # division rounds differently from * (1/3)
    "output() = sqrt(input());",
    "output() = convolve(mask, Reduce::SUM, [&] () -> float { return mask() * mask(); });",
    "output() = reduce(dom, Reduce::SUM, [&] () -> float { return input(dom) + 1; });",
    "if (input() > 3) output() = 1.0f;",
])
def test_unsupported_forms_are_refused(body):
    with pytest.raises(R.Unsupported):
        kernel(body)


def test_global_reduce_forms():
    assert R.classify_reduce(toks("{ return min(left, right); }"), ("left", "right")) == "MIN"
    assert R.classify_reduce(toks("{ return right + left; }"), ("left", "right")) == "SUM"
    with pytest.raises(R.Unsupported):
        R.classify_reduce(toks("{ return left - right; }"), ("left", "right"))


def _clang_ok():
    try:
        from hipacc_amd.emit import clang
        clang.lib()
        return os.path.exists(os.path.join(DSL_HEADERS, "hipacc.hpp"))
    except OSError:
        return False


@pytest.mark.skipif(not _clang_ok(), reason="libclang or the DSL headers are not available")
@pytest.mark.parametrize("name", sorted(p.stem for p in DSL.glob("*.cpp")))
def test_emitter_reproduces_committed_output(name):
    from hipacc_amd.emit import emit
    text = emit(str(DSL / f"{name}.cpp"), [DSL_HEADERS], fp_contract=False)
    committed = (DSL / "emitted" / f"{name}.cpp").read_text()
    assert text.replace(str(DSL) + "/", "") == committed


@pytest.mark.skipif(not _clang_ok(), reason="libclang or the DSL headers are not available")
def test_unrecognised_kernel_becomes_device_code(tmp_path):
    """A kernel() outside the library's operators is compiled as written
    (a device struct launched by hipaccLaunchUser); a kernel class with
    other member functions is refused."""
    from hipacc_amd.emit import EmitError, emit
    src = (DSL / "point_f32.cpp").read_text().replace(
        "output() = input() * 0.5f + 0.25f;", "output() = sqrtf(input());")
    p = tmp_path / "user.cpp"
    p.write_text(src)
    text = emit(str(p), [DSL_HEADERS])
    assert "struct Scale : hipacc_dev::Kernel<float>" in text
    assert "__device__ __forceinline__ void kernel() { output() = sqrtf(input()); }" in text
    assert "hipaccLaunchUser(scale, iscaled);" in text
    assert '#include "hipacc_hip_device.hpp"' in text
    bad = src.replace("    void kernel() { output() = sqrtf(input()); }",
                      "    float half(float v) { return v * 0.5f; }\n"
                      "    void kernel() { output() = half(input()); }")
    p.write_text(bad)
    with pytest.raises(EmitError, match="half"):
        emit(str(p), [DSL_HEADERS])


def test_fusible_chain_is_one_launch():
    """The Mixed fusion sample's chain, whose intermediate images nothing else
    names, is emitted as one hipaccLaunchFused; the point_f32 program's kernels,
    whose images are all written out, stay separate launches."""
    fused = (DSL / "emitted" / "fuse_mixed_u8.cpp").read_text()
    assert fused.count("hipaccLaunchFused(") == 1
    assert "hipaccLaunchLocal(" not in fused and "hipaccLaunchPoint(" not in fused
    assert fused.count("push_back(") == 8
    separate = (DSL / "emitted" / "point_f32.cpp").read_text()
    assert "hipaccLaunchFused(" not in separate
    assert separate.count("hipaccLaunchPoint(") == 3 and separate.count("hipaccLaunchLocal(") == 1


SAMPLES = Path("/root/reference/samples-public")


def _samples():
    return sorted(p.parent.parent.name for p in SAMPLES.glob("[0-6]_*/*/src/main.cpp"))


@pytest.mark.skipif(not _clang_ok() or not SAMPLES.exists(),
                    reason="libclang or the reference's samples are not available")
@pytest.mark.parametrize("name", _samples())
def test_emitter_takes_every_reference_sample(name):
    """Every sample program of the reference is emitted unchanged (the GPU
    test tests/test_gpu_samples.py runs them); a few structural facts of the
    rewriting are checked here."""
    from hipacc_amd.emit import emit
    src = next(SAMPLES.glob(f"*/{name}/src/main.cpp"))
    text = emit(str(src), [DSL_HEADERS, str(SAMPLES / "common")])
    assert "hipaccInitHIP(0);" in text
    main = text.split("HIPACC_CODEGEN int main", 1)[1]
    assert not re.search(r"\b(Image|Accessor|IterationSpace|BoundaryCondition|Pyramid)<", main)
    if name == "Gaussian_Laplacian_Pyramid":
        assert "HipaccPyramidTraversor hipaccTraversor;" in text
        assert text.count("hipaccTraversor.hipaccTraverse") == 2
        assert "HIPACC_INTERP_NN, &iter2" in text and "HIPACC_INTERP_LF" in text
        assert "hipaccLaunchLocal(acc1, iter1" in text     # the Gaussian stays a library call
    if name == "Game_of_Life":
        assert "dom.domain_at(0,0) = 0;" in text and "hipaccCopyMemory(out, in);" in text
    if name == "Motion_Interpolation":
        assert "hipaccWriteMemory<float4>(merged_vec, vecs);" in text
        assert "[&] { hipaccLaunchPoint(HIPACC_POINT_COPY, acc_vec, iter_vec);" in text
    if name == "Bilateral_Filter":                # the library's bilateral kernel
        assert "hipaccLaunchBilateral(acc, iter, mask.size_x" in text
        assert "hipaccLaunchUser" not in text
    if name.startswith("Kernel_Fusion_"):        # the whole chain in one launch
        assert text.count("hipaccLaunchFused(") == 1 and "hipaccLaunchUser" not in text
