"""HIP emitter for Hipacc DSL programs: the `-emit-hip` target.

The reference's source-to-source compiler rewrites a DSL program's host code
into calls of its CUDA runtime and generates one kernel per Kernel class
(lib/Rewrite/CreateHostStrings.cpp:650-1198, lib/Backend/CUDA.cpp:38-190).
Here the kernels already exist in libhipacc_gfx950.so, so the emitter only
has to recognise which operator each Kernel class is and rewrite main():

  Image<T> img(w, h[, host])          -> HipaccImageHip<T> img = hipaccCreateMemory<T>(host, w, h)
  Mask<T> m(coef)                     -> HipaccMask<T> m(sx, sy, &coef[0][0])
  Domain d(sx, sy) / Domain d(mask)   -> HipaccMask<float> d(sx, sy) / a copy with holes
  BoundaryCondition<T> b(img, m, MODE[, v])   (recorded; applied at the launch)
  Accessor<T> a(b | img[, w, h, x, y])        -> HipaccAccessor<T> a = hipaccMakeAccessor(img, ...)
  IterationSpace<T> is(img[, w, h, x, y])     -> HipaccAccessor<T> is = hipaccMakeAccessor(img, ...)
  K k(is, a, m, ...)                  (recorded)
  k.execute()                         -> hipaccLaunchLocal / hipaccLaunchPoint / hipaccApplyReduction
  k.reduced_data()                    -> the reduction's value
  img.data()                          -> hipaccReadMemory<T>(img)

Kernel classes are parsed with libclang (hipacc_amd/emit/clang.py) and
matched by hipacc_amd/emit/recognise.py; a kernel outside the recognised
forms stops the emitter with the reason.  Usage:

    python -m hipacc_amd.emit prog.cpp -o prog_hip.cpp -I <dsl include dir> [--no-fp-contract]
"""
from __future__ import annotations

import re
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

from . import clang as C
from .recognise import (BilateralOp, GlobalOp, LocalOp, PointOp, Unsupported,
                        classify_bilateral, classify_kernel, classify_reduce)


class EmitError(RuntimeError):
    pass


@dataclass
class UserOp:
    """A kernel() the library has no operator for: compiled as it is, as
    device code against hipacc_hip_device.hpp."""
    body: str
    reason: str
    kind: str = "user"


@dataclass
class KernelClass:
    name: str
    out_type: str
    fields: Dict[str, Tuple[str, str]] = field(default_factory=dict)  # name -> (kind, elem type)
    params: List[str] = field(default_factory=list)                  # constructor parameters
    param_field: Dict[str, str] = field(default_factory=dict)         # parameter -> member
    op: object = None
    global_reduce: Optional[str] = None   # SUM MIN MAX PROD, or USER
    bin_type: str = ""                    # Kernel<T, TB>: TB
    reduce_text: str = ""                 # reduce(l, r) as written
    binning_text: str = ""                # binning(x, y, pixel) as written
    bin_reduce: Optional[str] = None      # reduce() of a binning class: SUM MIN MAX PROD
    body_text: str = ""                   # kernel() body as written
    extent: Tuple[int, int] = (0, 0)


@dataclass
class AccInfo:
    """An Accessor / IterationSpace of main(): the image expression it reads
    (an image, or a pyramid level `p(i)`), its boundary handling, its
    interpolation (None, or NN / LF / B5 / CF / L3), whether it crops, and
    the name it was constructed from (image or boundary condition)."""
    img: str
    mode: str
    value: str
    interp: Optional[str]
    crop: bool
    src: str


_TPL = re.compile(r"^(?:const\s+)?(?:hipacc::)?(\w+)\s*(?:<\s*(.+?)\s*>)?\s*&?$")


def _tpl(type_spelling: str) -> Tuple[str, str]:
    m = _TPL.match(type_spelling.strip())
    if not m:
        return "", ""
    return m.group(1), (m.group(2) or "")


def _args(tokens: Sequence[str], start: int) -> Tuple[List[str], int]:
    """Top-level comma-separated arguments of the parenthesis at tokens[start]."""
    assert tokens[start] == "("
    depth, cur, out = 0, [], []
    for j in range(start, len(tokens)):
        t = tokens[j]
        if t in "([{":
            depth += 1
            if depth == 1:
                continue
        elif t in ")]}":
            depth -= 1
            if depth == 0:
                if cur:
                    out.append(_join(cur))
                return out, j
        if depth == 1 and t == ",":
            out.append(_join(cur))
            cur = []
        else:
            cur.append(t)
    raise EmitError("unbalanced parentheses")


def _member_inits(toks: Sequence[str], params: Sequence[str]) -> Dict[str, str]:
    """parameter -> member of a constructor's `member(parameter)` initialisers."""
    depth, start = 0, None
    for j, t in enumerate(toks):
        if t == "(":
            depth += 1
        elif t == ")":
            depth -= 1
        elif t == ":" and depth == 0:
            start = j + 1
            break
    out: Dict[str, str] = {}
    if start is None:
        return out
    j = start
    while j + 3 < len(toks) and toks[j] != "{":
        if toks[j + 1] == "(" and toks[j + 3] == ")" and toks[j + 2] in params:
            out[toks[j + 2]] = toks[j]
        j += 1
    return out


def _join(toks: Sequence[str]) -> str:
    s = ""
    for t in toks:
        if s and (re.match(r"\w", t) and re.match(r".*\w$", s)):
            s += " "
        s += t
    return s


class Emitter:
    def __init__(self, path: str, include_dirs: Sequence[str] = (), defines: Sequence[str] = (),
                 fp_contract: bool = True):
        self.path = path
        self.fp_contract = fp_contract
        args = ["-x", "c++", "-std=c++14"]
        rd = C.resource_dir()
        if rd:
            args += ["-resource-dir", rd]
        args += [f"-I{d}" for d in include_dirs] + [f"-D{d}" for d in defines]
        self.tu = C.TranslationUnit(path, args, C.DETAILED_PREPROCESSING)
        with open(path, "rb") as f:
            self.src = f.read()
        main_errors = [e for e in self.tu.errors if e.startswith(path)]
        if main_errors:
            raise EmitError("the program does not parse:\n" + "\n".join(main_errors))
        self.macros = self._object_macros()
        self.kernels: Dict[str, KernelClass] = {}
        self.field_types: Dict[Tuple[str, str], str] = {}
        self.var_types: Dict[str, str] = {}
        self.edits: List[Tuple[int, int, str]] = []

    def _object_macros(self) -> Dict[str, List[str]]:
        """The program's object-like macros (name -> replacement tokens), for
        kernel bodies written with them (`TYPE pixel = in();`)."""
        out: Dict[str, List[str]] = {}
        for c in self.tu.cursor.children():
            if c.kind != "macro definition" or not c.in_main_file():
                continue
            b, _ = c.extent()
            name = c.spelling
            if self.src[b + len(name):b + len(name) + 1] == b"(":
                continue                          # function-like
            out[name] = c.tokens()[1:]
        return out

    def _expand(self, toks: Sequence[str], depth: int = 0) -> List[str]:
        out: List[str] = []
        for t in toks:
            rep = self.macros.get(t)
            if rep is None or depth >= 8:
                out.append(t)
            else:
                out.extend(self._expand(rep, depth + 1))
        return out

    # ---- kernel classes ---------------------------------------------------------
    def _kernel_class(self, c: C.Cursor) -> Optional[KernelClass]:
        base = [ch for ch in c.children() if ch.kind == "C++ base class specifier"]
        if len(base) != 1:
            return None
        name, targ = _tpl(base[0].type)
        if name != "Kernel":
            return None
        targs = [t.strip() for t in targ.split(",")]
        k = KernelClass(c.spelling, targs[0], extent=c.extent())
        k.bin_type = targs[1] if len(targs) > 1 else ""

        def method_text(m: C.Cursor) -> str:
            b0, b1 = m.extent()
            return self.src[b0:b1].decode()
        kernel_body = reduce_body = None
        kernel_text = "" 
        reduce_params: Tuple[str, str] = ("", "")
        for ch in c.children():
            if ch.kind == "FieldDecl":
                kind, elem = _tpl(ch.type)
                k.fields[ch.spelling] = (kind, elem)
                self.field_types[(k.name, ch.spelling)] = ch.type
            elif ch.kind == "CXXConstructor":
                sub = ch.children()
                k.params = [p.spelling for p in sub if p.kind == "ParmDecl"]
                k.param_field.update(_member_inits(ch.tokens(), k.params))
            elif ch.kind == "CXXMethod" and ch.spelling == "kernel":
                body = [s for s in ch.children() if s.kind == "CompoundStmt"]
                kernel_body = self._expand(body[0].tokens()) if body else None
                if body:
                    b0, b1 = body[0].extent()
                    kernel_text = self.src[b0:b1].decode()
            elif ch.kind == "CXXMethod" and ch.spelling not in ("reduce", "binning"):
                raise EmitError(f"kernel class {k.name}: member function {ch.spelling}() "
                                "is not emitted")
            elif ch.kind == "CXXMethod" and ch.spelling == "reduce":
                ps = [p.spelling for p in ch.children() if p.kind == "ParmDecl"]
                body = [s for s in ch.children() if s.kind == "CompoundStmt"]
                if len(ps) == 2 and body:
                    reduce_body, reduce_params = self._expand(body[0].tokens()), (ps[0], ps[1])
                    k.reduce_text = method_text(ch)
            elif ch.kind == "CXXMethod" and ch.spelling == "binning":
                k.binning_text = method_text(ch)
        if kernel_body is None:
            raise EmitError(f"kernel class {k.name}: no kernel() body")
        k.body_text = kernel_text
        acc = [n for n, (kd, _) in k.fields.items() if kd == "Accessor"]
        msk = [n for n, (kd, _) in k.fields.items() if kd == "Mask"]
        dom = [n for n, (kd, _) in k.fields.items() if kd == "Domain"]
        if reduce_body is not None and not k.binning_text:
            try:
                k.global_reduce = classify_reduce(reduce_body, reduce_params)
            except Unsupported:
                k.global_reduce = "USER"
        if k.binning_text and not k.reduce_text:
            raise EmitError(f"kernel class {k.name}: binning() without reduce()")
        if k.binning_text and reduce_body is not None:
            try:
                k.bin_reduce = classify_reduce(reduce_body, reduce_params)
            except Unsupported:
                k.bin_reduce = None
        try:
            k.op = classify_kernel(kernel_body, acc, msk, dom)
        except Unsupported as e:
            k.op = UserOp(kernel_text, str(e))
            scalars = [n for n, (kd, _) in k.fields.items()
                       if kd not in ("Accessor", "Mask", "Domain")]
            try:
                op = classify_bilateral(kernel_body, acc, msk, dom, scalars)
                pix = k.fields[op.accessor][1]
                if (op.round and pix in ("uchar", "unsigned char") and
                        k.out_type in ("uchar", "unsigned char")) or \
                        (not op.round and pix == "float" and k.out_type == "float"):
                    k.op = op
            except Unsupported:
                pass
        return k

    # ---- main() -----------------------------------------------------------------------
    def _collect_var_types(self, c: C.Cursor) -> None:
        for ch in c.walk():
            if ch.kind == "VarDecl" and ch.in_main_file():
                self.var_types[ch.spelling] = ch.type

    def run(self) -> str:
        main = None
        for c in self.tu.cursor.children():
            if not c.in_main_file():
                continue
            if c.kind in ("ClassDecl", "StructDecl"):
                k = self._kernel_class(c)
                if k:
                    self.kernels[k.name] = k
            elif c.kind == "FunctionDecl" and c.spelling == "main":
                if any(ch.kind == "CompoundStmt" for ch in c.children()):
                    main = c
            elif c.kind == "VarDecl":
                self.var_types[c.spelling] = c.type
        if main is None:
            raise EmitError("no main() definition")
        self._collect_var_types(main)
        body = [ch for ch in main.children() if ch.kind == "CompoundStmt"][0]
        mb, me = main.extent()
        self.main_text = self.src[mb:me].decode()
        # A kernel the library would run as an operator but whose instance
        # reads through an interpolating accessor is compiled as device code
        # instead, and the accessor members some instance binds to an
        # interpolating accessor are declared as such; main() is walked
        # again until both are settled.
        self.interp_fields: Dict[str, set] = {}
        for _ in range(3):
            self._walk_main(body)
            forced = {self.instances[i][0] for i in self.interp_insts}
            forced = {n for n in forced if not isinstance(self.kernels[n].op, UserOp)}
            fields = self._interp_fields()
            if not forced and fields == self.interp_fields:
                break
            self.interp_fields = fields
            for n in forced:
                k = self.kernels[n]
                k.op = UserOp(k.body_text, "reads through an interpolating accessor")
        for k in self.kernels.values():
            b, e = k.extent
            if self.src[e:e + 1] == b";":
                e += 1
            text = f"// Kernel class {k.name}: {_describe(k)}"
            if isinstance(k.op, UserOp):
                text += "\n" + self._user_struct(k)
            scalars = "".join(f"    {ty} {n};\n" for n, ty in self._scalar_members(k))
            if k.binning_text:
                text += (f"\nstruct {k.name}_bins : hipacc_dev::Binning<{k.bin_type}> {{\n"
                         f"{scalars}    __device__ {k.binning_text}\n"
                         f"    __device__ {k.reduce_text}\n}};")
            elif k.global_reduce == "USER":
                text += (f"\nstruct {k.name}_reduce {{\n{scalars}"
                         f"    __device__ {k.reduce_text}\n}};")
            self.edits.append((b, e, text))
        self.user = any(isinstance(k.op, UserOp) or k.binning_text or k.global_reduce == "USER"
                        for k in self.kernels.values())
        for chain in self._chains():
            if len(chain) == 1:
                inst, b, e = chain[0]
                self.edits.append((b, e, self._launch(inst)))
                continue
            for inst, b, e in chain[:-1]:
                self.edits.append((b, e, f"// {inst}.execute(): fused into the launch of "
                                         f"{chain[-1][0]}"))
            inst, b, e = chain[-1]
            self.edits.append((b, e, self._fused_launch([c[0] for c in chain])))
        self._inline_calls(main)
        b0 = body.extent()[0]
        init = "\n    hipaccInitHIP(0);" + ("\n    hipaccSetTiming(true);" if self.timing else "")
        self.edits.append((b0 + 1, b0 + 1, init))
        return self._apply()

    def _walk_main(self, body: C.Cursor) -> None:
        self.edits = []
        self.images: Dict[str, str] = {}
        self.pyramids: Dict[str, str] = {}
        self.bcs: Dict[str, Tuple[str, str, str]] = {}
        self.accessors: Dict[str, AccInfo] = {}
        self.instances: Dict[str, Tuple[str, List[str]]] = {}
        self.inst_accs: Dict[str, Dict[str, AccInfo]] = {}
        self.interp_insts: List[str] = []
        self.masks: Dict[str, str] = {}
        self.domains: set = set()
        self.timing = b"hipacc_last_kernel_timing" in self.src
        self.launches: List[Tuple[str, int, int]] = []
        self.launch_scope: Dict[int, Tuple[int, int]] = {}
        self.image_args: Dict[str, List[str]] = {}
        self.traversor = False
        for st in body.children():
            self._visit(st, body.extent())

    def _visit(self, st: C.Cursor, scope: Tuple[int, int]) -> None:
        """A statement of main() (or of a block, loop or lambda inside it)."""
        if st.kind == "DeclStmt":
            self._statement(st)
            return
        if st.kind == "CallExpr" and self._call(st, scope):
            return
        for ch in st.children():
            self._visit(ch, ch.extent() if ch.kind == "CompoundStmt" else scope)

    def _statement(self, st: C.Cursor) -> None:
        """A declaration statement, one or more declarators of one type."""
        b, e = st.extent()
        if self.src[e:e + 1] == b";":
            e += 1
        reps = []
        for v in (ch for ch in st.children() if ch.kind == "VarDecl"):
            toks = v.tokens()
            name, ty = v.spelling, v.type
            kind, targ = _tpl(ty)
            i = toks.index(name) if name in toks else -1
            args = (_args(toks, i + 1)[0] if 0 <= i < len(toks) - 1 and toks[i + 1] == "("
                    else [])
            reps.append((name, self._declaration(name, kind, targ, args)))
        if all(r is None for _, r in reps):
            return
        if any(r is None for _, r in reps):
            raise EmitError("a declaration mixes DSL objects and other variables: "
                            + ", ".join(n for n, _ in reps))
        self.edits.append((b, e, "\n    ".join(r for _, r in reps)))

    def _call(self, st: C.Cursor, scope: Tuple[int, int]) -> bool:
        """A call statement the emitter rewrites; False to look inside it."""
        toks = st.tokens()
        text = "".join(toks).rstrip(";")
        b, e = st.extent()
        end = e + 1 if self.src[e:e + 1] == b";" else e
        m = re.match(r"^(\w+)\.execute\(\)$", text)
        if m and m.group(1) in self.instances:
            self.launches.append((m.group(1), b, end))
            self.launch_scope[b] = scope
            return True
        if toks and toks[0] == "traverse":
            # traverse(...) -> the runtime's traversor (lib/Rewrite/Rewrite.cpp:
            # 2083-2100 declares it before the first call)
            rep = "hipaccTraversor.hipaccTraverse"
            if not self.traversor:
                rep = "HipaccPyramidTraversor hipaccTraversor;\n    " + rep
                self.traversor = True
            self.edits.append((b, b + len("traverse"), rep))
            for ch in st.children():
                self._visit(ch, scope)
            return True
        m = re.match(r"^(\w+)=(\w+)$", text)
        if m and m.group(1) in self.images:
            dst, src = m.group(1), m.group(2)
            if src in self.images:      # Image = Image: a device copy
                rep = f"hipaccCopyMemory({src}, {dst});"
            else:                        # Image = host pointer: an upload
                rep = f"hipaccWriteMemory<{self.images[dst]}>({dst}, {src});"
            self.edits.append((b, end, rep))
            return True
        m = re.match(r"^(\w+)\((.+)\)=(.+)$", text)
        if m and m.group(1) in self.masks:
            what = "domain_at" if m.group(1) in self.domains else "coef_at"
            self.edits.append((b, end, f"{m.group(1)}.{what}({m.group(2)}) = {m.group(3)};"))
            return True
        return False

    def _img_type(self, expr: str) -> Optional[str]:
        """The pixel type of an image expression: an image, or a pyramid level."""
        if expr in self.images:
            return self.images[expr]
        m = re.match(r"^(\w+)\((.+)\)$", expr)
        if m and m.group(1) in self.pyramids:
            return self.pyramids[m.group(1)]
        return None

    def _declaration(self, name: str, kind: str, targ: str, args: List[str]) -> Optional[str]:
        if kind == "Image":
            self.images[name] = targ
            self.image_args[name] = args[:2]
            host = args[2] if len(args) > 2 else "nullptr"
            if len(args) < 2:
                raise EmitError(f"Image {name}: expected (width, height[, data])")
            return f"HipaccImageHip<{targ}> {name} = hipaccCreateMemory<{targ}>({host}, {args[0]}, {args[1]});"
        if kind == "Pyramid":
            if len(args) != 2 or args[0] not in self.images:
                raise EmitError(f"Pyramid {name}: expected (image, depth)")
            self.pyramids[name] = targ
            return (f"HipaccPyramidHip<{targ}> {name} = "
                    f"hipaccCreatePyramid<{targ}>({args[0]}, {args[1]});")
        if kind == "Mask":
            if len(args) != 1:
                raise EmitError(f"Mask {name}: expected one coefficient array")
            dims = re.findall(r"\[(\d+)\]", self.var_types.get(args[0], ""))
            if len(dims) != 2:
                raise EmitError(f"Mask {name}: the size of {args[0]} is not a constant 2D array")
            self.masks[name] = targ
            sy, sx = dims
            return f"HipaccMask<{targ}> {name}({sx}, {sy}, &{args[0]}[0][0]);"
        if kind == "Domain":
            self.domains.add(name)
            if len(args) == 2:
                self.masks[name] = "float"
                return f"HipaccMask<float> {name}({args[0]}, {args[1]});"
            if len(args) == 1 and args[0] in self.masks:
                self.masks[name] = self.masks[args[0]]
                return f"auto {name} = {args[0]};\n    {name}.domain_from_coef();"
            dims = re.findall(r"\[(\d+)\]", self.var_types.get(args[0], "")) if args else []
            if len(args) == 1 and len(dims) == 2:
                # Domain(const uchar (&)[sy][sx]): the array's nonzero entries
                self.masks[name] = "float"
                sy, sx = dims
                return (f"HipaccMask<float> {name}({sx}, {sy});\n"
                        f"    {name}.domain_write(&{args[0]}[0][0]);")
            raise EmitError(f"Domain {name}: expected (size_x, size_y), (mask) or a 2D array")
        if kind == "BoundaryCondition":
            mode = [a for a in args if re.match(r"^(?:hipacc::)?Boundary::\w+$", a)]
            if not args or len(mode) != 1:
                raise EmitError(f"BoundaryCondition {name}: no Boundary mode")
            mi = args.index(mode[0])
            value = args[mi + 1] if len(args) > mi + 1 else "0"
            self.bcs[name] = (args[0], mode[0].split("::")[-1], value)
            return f"// BoundaryCondition {name}: {mode[0].split('::')[-1]} of {args[0]}"
        if kind in ("Accessor", "IterationSpace"):
            if not args:
                raise EmitError(f"{kind} {name}: no image")
            ip = [a for a in args if re.match(r"^(?:hipacc::)?Interpolate::\w+$", a)]
            args = [a for a in args if a not in ip]
            interp = ip[0].split("::")[-1] if ip else None
            if interp == "NO":
                interp = None
            if interp and kind != "Accessor":
                raise EmitError(f"{kind} {name}: an iteration space does not interpolate")
            src = args[0]
            if src in self.bcs:
                img, mode, value = self.bcs[src]
            else:
                img, mode, value = src, "UNDEFINED", "0"
            if self._img_type(img) is None:
                raise EmitError(f"{kind} {name}: unknown image {img}")
            if len(args) not in (1, 5):
                raise EmitError(f"{kind} {name}: expected (image) or (image, w, h, x, y)")
            self.accessors[name] = AccInfo(img, mode, value, interp, len(args) != 1, src)
            crop = "" if len(args) == 1 else ", " + ", ".join(args[1:5])
            return f"HipaccAccessor<{targ}> {name} = hipaccMakeAccessor({img}{crop});"
        if kind in self.kernels:
            self.instances[name] = (kind, args)
            self.inst_accs[name] = {a: self.accessors[a] for a in args if a in self.accessors}
            if any(a.interp for a in self.inst_accs[name].values()):
                self.interp_insts.append(name)
            if isinstance(self.kernels[kind].op, UserOp):
                return self._user_instance(name)
            return f"// {kind} {name}({', '.join(args)}): launched at {name}.execute()"
        return None

    def _interp_fields(self) -> Dict[str, set]:
        """Kernel class -> its accessor members bound to an interpolating
        accessor by some instance."""
        out: Dict[str, set] = {}
        for inst in self.interp_insts:
            k, bind, _ = self._binding(inst)
            for f, (kind, _) in k.fields.items():
                if kind == "Accessor" and f in bind and self._acc(inst, bind[f]).interp:
                    out.setdefault(k.name, set()).add(f)
        return out

    def _acc(self, inst: str, name: str) -> AccInfo:
        """Accessor `name` as it was when instance `inst` was declared."""
        return self.inst_accs[inst][name]

    def _binding(self, inst: str) -> Tuple[KernelClass, Dict[str, str], str]:
        kname, args = self.instances[inst]
        k = self.kernels[kname]
        if len(args) != len(k.params):
            raise EmitError(f"{inst}: {len(args)} arguments for {len(k.params)} parameters")
        bind = {}
        for p, a in zip(k.params, args):
            if p in k.param_field:
                bind[k.param_field[p]] = a
        iter_name = args[0] if args else ""
        if iter_name not in self.inst_accs.get(inst, {}):
            raise EmitError(f"{inst}: the first argument is not an IterationSpace")
        return k, bind, iter_name

    def _launch(self, inst: str) -> str:
        k, bind, it = self._binding(inst)
        op = k.op
        if isinstance(op, UserOp):
            return f"hipaccLaunchUser({inst}, {it});"
        if isinstance(op, BilateralOp):
            a = self._acc(inst, bind[op.accessor])
            sp = f"{inst}_spatial"
            return ("{\n        "
                    f"auto {sp} = hipaccBilateralSpatial({bind[op.mask]}, {bind[op.domain]});\n"
                    f"        hipaccLaunchBilateral({bind[op.accessor]}, {it}, "
                    f"{bind[op.mask]}.size_x, 0.0f, (float)({bind[op.sigma_r]}), {sp}.data(), "
                    f"HipaccBoundary::{a.mode}, {a.value}, {'true' if op.round else 'false'});"
                    "\n    }")
        if isinstance(op, PointOp):
            ins = [bind[a] for a in op.inputs]
            if op.op in ("COPY", "AFFINE"):
                t0 = self._img_type(self._acc(inst, ins[0]).img)
                ab = ("" if op.op == "COPY" else
                      f", static_cast<const HipaccAccessor<{t0}>*>(nullptr), {op.alpha}, {op.beta}")
                return f"hipaccLaunchPoint(HIPACC_POINT_{op.op}, {ins[0]}, {it}{ab});"
            return f"hipaccLaunchPoint(HIPACC_POINT_{op.op}, {ins[0]}, {it}, &{ins[1]});"
        assert isinstance(op, LocalOp)
        acc = bind[op.accessor]
        lines, o, m = self._local_setup(inst, op, bind)
        lines.append(f"hipaccLaunchLocal({acc}, {it}, {m}, {o});")
        return "{\n        " + "\n        ".join(lines) + "\n    }"

    def _local_setup(self, inst: str, op: LocalOp, bind: Dict[str, str]):
        """Declarations of a local launch's options and mask."""
        a = self._acc(inst, bind[op.accessor])
        mode, value = a.mode, a.value
        mtype = "float" if op.acc_type in ("float", "double") else "int"
        o = f"{inst}_opt"
        lines = [f"HipaccLocalOptions {o};",
                 f"{o}.reduce = HipaccReduce::{op.reduce};",
                 f"{o}.boundary = HipaccBoundary::{mode};"]
        if mode == "CONSTANT":
            lines.append(f"{o}.const_val = {value};")
        lines.append(f"{o}.use_weights = {'true' if op.mask else 'false'};")
        if op.round:
            lines.append(f"{o}.round = true;")
        if op.post_div:
            lines.append(f"{o}.post_div = {op.post_div};")
        lines.append(f"{o}.fp_contract = {'true' if self.fp_contract else 'false'};")
        m = f"{inst}_mask"
        if op.mask and op.domain:
            lines.append(f"auto {m} = hipaccMaskCast<{mtype}>({bind[op.mask]});")
            lines.append(f"{m}.domain = {bind[op.domain]}.domain;")
        else:
            lines.append(f"auto {m} = hipaccMaskCast<{mtype}>({bind[op.mask or op.domain]});")
        return lines, o, m

    # ---- user kernels -----------------------------------------------------------------
    def _user_struct(self, k: KernelClass) -> str:
        members, pos = [], ["it_x = x_;", "it_y = y_;"]
        for name, (kind, elem) in k.fields.items():
            if kind == "Accessor":
                ip = ", true" if name in self.interp_fields.get(k.name, ()) else ""
                members.append(f"hipacc_dev::Accessor<{elem}{ip}> {name};")
                pos.append(f"{name}.px = x_; {name}.py = y_;")
            elif kind == "Mask":
                members.append(f"hipacc_dev::Mask<{elem}> {name};")
            elif kind == "Domain":
                members.append(f"hipacc_dev::Domain {name};")
            elif kind in ("IterationSpace", "Image", "BoundaryCondition"):
                raise EmitError(f"kernel class {k.name}: member {name} of type {kind}")
            else:
                # scalar members are set from the host after construction
                members.append(f"{dict(self._scalar_members(k))[name]} {name};")
        lines = [f"struct {k.name} : hipacc_dev::Kernel<{k.out_type}> {{"]
        lines += ["    " + m for m in members]
        lines.append("    __device__ void set_pos(int x_, int y_) { " + " ".join(pos) + " }")
        lines.append("    __device__ __forceinline__ void kernel() " + k.op.body)
        lines.append("};")
        return "\n".join(lines)

    def _user_instance(self, inst: str) -> str:
        k, bind, it = self._binding(inst)
        out = [f"{k.name} {inst};"]
        for name, (kind, elem) in k.fields.items():
            if name not in bind:
                raise EmitError(f"{inst}: member {name} is not set by the constructor")
            arg = bind[name]
            if kind == "Accessor":
                a = self._acc(inst, arg)
                ip = f", HIPACC_INTERP_{a.interp}, &{it}" if a.interp else ""
                tpl = "<true>" if name in self.interp_fields.get(k.name, ()) else ""
                out.append(f"{inst}.{name} = hipaccDevAccessor{tpl}({arg}, "
                           f"HipaccBoundary::{a.mode}, {a.value}{ip});")
            elif kind == "Mask":
                out.append(f"{inst}.{name} = hipaccDevMask({arg});")
            elif kind == "Domain":
                out.append(f"{inst}.{name} = hipaccDevDomain({arg});")
            else:
                out.append(f"{inst}.{name} = {arg};")
        return "\n    ".join(out)

    # ---- fusion ---------------------------------------------------------------------
    _FUSED_TYPES = {"uchar", "unsigned char", "char", "signed char", "short", "ushort",
                    "unsigned short", "int", "float"}

    def _chain_input(self, inst: str) -> Optional[str]:
        """The accessor a fusible kernel reads its chain value through."""
        k, bind, _ = self._binding(inst)
        if (k.global_reduce or k.binning_text or isinstance(k.op, UserOp) or
                k.out_type not in self._FUSED_TYPES):
            return None
        if isinstance(k.op, LocalOp):
            return bind[k.op.accessor]
        if not isinstance(k.op, PointOp):
            return None
        if k.op.op in ("COPY", "AFFINE", "ADD", "SUB", "MUL"):
            return bind[k.op.inputs[0]]
        return None

    def _fusible(self, prev: str, nxt: str) -> bool:
        """Whether kernel nxt can take kernel prev's output image without the
        image: nxt reads it through its chain input only, nothing else in
        main() names it, and every image involved has the same size."""
        a_prev, a_next = self._chain_input(prev), self._chain_input(nxt)
        if a_prev is None or a_next is None:
            return False
        _, _, it = self._binding(prev)
        k, bind, it_next = self._binding(nxt)
        acc_it, acc_next = self._acc(prev, it), self._acc(nxt, a_next)
        x = acc_it.img
        if acc_next.img != x or acc_it.crop or acc_next.crop:
            return False
        if self.images.get(x) not in self._FUSED_TYPES:
            return False
        for other in (k.op.inputs[1:] if isinstance(k.op, PointOp) else []):
            if self._acc(nxt, bind[other]).img == x:
                return False
        imgs = [self._acc(prev, a_prev).img, x, self._acc(nxt, it_next).img]
        if any(i not in self.image_args for i in imgs):
            return False
        if len({tuple(self.image_args[i]) for i in imgs}) != 1:
            return False
        # the image, its iteration space and the accessor (and boundary
        # condition) reading it are the only names of it in main()
        decls = 1 + sum(1 for a in self.accessors.values() if a.src == x)
        decls += sum(1 for n, (img, _, _) in self.bcs.items() if img == x)
        if len(re.findall(r"\b%s\b" % re.escape(x), self.main_text)) != decls:
            return False
        users = [i for i, (_, args) in self.instances.items() if a_next in args or it in args]
        return sorted(users) == sorted({prev, nxt})

    def _chains(self) -> List[List[Tuple[str, int, int]]]:
        chains: List[List[Tuple[str, int, int]]] = []
        for launch in self.launches:
            if chains and self._extends(chains[-1], launch):
                chains[-1].append(launch)
            else:
                chains.append([launch])
        return chains

    def _extends(self, chain, launch: Tuple[str, int, int]) -> bool:
        inst = launch[0]
        if self.launch_scope[chain[-1][1]] != self.launch_scope[launch[1]]:
            return False
        if len(chain) >= 8 or not self._fusible(chain[-1][0], inst):
            return False
        locals_ = sum(isinstance(self._binding(c[0])[0].op, LocalOp) for c in chain)
        return locals_ + isinstance(self._binding(inst)[0].op, LocalOp) <= 2

    def _fused_launch(self, insts: List[str]) -> str:
        name = f"{insts[-1]}_chain"
        lines = [f"std::vector<HipaccFusedStage> {name};"]
        for inst in insts:
            k, bind, _ = self._binding(inst)
            op = k.op
            if isinstance(op, LocalOp):
                ll, o, m = self._local_setup(inst, op, bind)
                lines += ll
                lines.append(f"{name}.push_back(hipaccFusedLocal<{k.out_type}>({m}, {o}));")
            elif op.op in ("COPY", "AFFINE"):
                lines.append(f"{name}.push_back(hipaccFusedPoint<{k.out_type}>(HIPACC_POINT_{op.op}, "
                             f"{op.alpha}, {op.beta}));")
            else:
                lines.append(f"{name}.push_back(hipaccFusedPoint2<{k.out_type}>("
                             f"HIPACC_POINT_{op.op}, {bind[op.inputs[1]]}));")
        first = self._chain_input(insts[0])
        last_it = self._binding(insts[-1])[2]
        lines.append(f"hipaccLaunchFused({first}, {last_it}, {name});")
        return ("{\n        // fused: " + " -> ".join(insts) + "\n        " +
                "\n        ".join(lines) + "\n    }")

    def _inline_calls(self, main: C.Cursor) -> None:
        """img.data() and k.reduced_data() anywhere in main()."""
        b, e = main.extent()
        text = self.src[b:e].decode()
        for m in re.finditer(r"\b(\w+)\s*\.\s*binned_data\s*\(([^()]*)\)", text):
            name = m.group(1)
            if name not in self.instances:
                continue
            k, _, it = self._binding(name)
            if not k.binning_text:
                raise EmitError(f"{name}.binned_data(): {k.name} has no binning()")
            red = (f", HIPACC_REDUCE_{k.bin_reduce}" if k.bin_reduce in ("SUM", "MIN", "MAX")
                   else "")
            rep = self._run_first(name, f"hipaccBinUser<{k.name}_bins, {k.bin_type}{red}>("
                                        f"{self._functor(name, '_bins')}, {it}, "
                                        f"{m.group(2).strip()})")
            s0 = b + len(text[:m.start()].encode())
            s1 = b + len(text[:m.end()].encode())
            if not any(x0 <= s0 < x1 for x0, x1, _ in self.edits):
                self.edits.append((s0, s1, rep))
        for m in re.finditer(r"\b(\w+)\s*\.\s*(data|reduced_data)\s*\(\s*\)", text):
            name, what = m.group(1), m.group(2)
            if what == "data" and name in self.images:
                rep = f"hipaccReadMemory<{self.images[name]}>({name})"
            elif what == "reduced_data" and name in self.instances:
                k, _, it = self._binding(name)
                if k.global_reduce == "USER":
                    rep = f"hipaccReduceUser({self._functor(name, '_reduce')}, {it})"
                elif k.global_reduce:
                    rep = f"hipaccApplyReduction<{k.out_type}>({it}, HipaccReduce::{k.global_reduce})"
                else:
                    raise EmitError(f"{name}.reduced_data(): {k.name} has no reduce()")
                rep = self._run_first(name, rep)
            else:
                continue
            s0 = b + len(text[:m.start()].encode())
            s1 = b + len(text[:m.end()].encode())
            if not any(x0 <= s0 < x1 for x0, x1, _ in self.edits):
                self.edits.append((s0, s1, rep))

    def _scalar_members(self, k: KernelClass) -> List[Tuple[str, str]]:
        """(name, type) of a kernel class's members other than accessors,
        masks and domains, without const or reference."""
        out = []
        for name, (kind, _) in k.fields.items():
            if kind not in ("Accessor", "Mask", "Domain", "IterationSpace", "Image",
                            "BoundaryCondition"):
                ty = re.sub(r"^const\s+|\s*&$", "", self.field_types[(k.name, name)].strip())
                out.append((name, ty))
        return out

    def _functor(self, inst: str, suffix: str) -> str:
        """The binning / reduce struct of an instance, its scalar members set
        from the instance's constructor arguments."""
        k, bind, _ = self._binding(inst)
        sets = [f"f.{n} = {bind[n]};" for n, _ in self._scalar_members(k) if n in bind]
        if not sets:
            return f"{k.name}{suffix}{{}}"
        return f"[&] {{ {k.name}{suffix} f; {' '.join(sets)} return f; }}()"

    def _run_first(self, inst: str, expr: str) -> str:
        """reduced_data() / binned_data() of a kernel never executed: the DSL
        executes it first (dsl/kernel.hpp:121-125,162-166)."""
        if any(l[0] == inst for l in self.launches):
            return expr
        return f"[&] {{ {self._launch(inst)} return {expr}; }}()"

    def _apply(self) -> str:
        out = self.src
        for b, e, rep in sorted(self.edits, key=lambda t: t[0], reverse=True):
            # statements nested deeper than main()'s body: their continuation
            # lines follow the statement's own indentation
            line = out[out.rfind(b"\n", 0, b) + 1:b]
            if len(line) > 4 and not line.strip():
                rep = rep.replace("\n    ", "\n" + line.decode())
            out = out[:b] + rep.encode() + out[e:]
        text = out.decode()
        # 4-wide pixel types need HIP's vector types (a hipcc build).
        vec = re.search(r"\b(u?char|u?short|u?int|u?long|float|double)4\b", text)
        prologue = (("#include \"hipacc_hip_device.hpp\"\n" if self.user else
                     "#include \"hipacc_hip.hpp\"\n") +
                    ("#include \"hipacc_hip_vec.hpp\"\n" if vec and not self.user else "") +
                    "// DSL names kept by the emitted host code\n"
                    "typedef unsigned char uchar;\n"
                    "typedef unsigned short ushort;\n"
                    "typedef unsigned int uint;\n"
                    "#ifndef HIPACC_CODEGEN\n#define HIPACC_CODEGEN\n#endif\n"
                    "namespace hipacc { namespace math { using std::min; using std::max; } }" +
                    ("\nnamespace hipacc { using Reduce = hipacc_dev::Reduce; }" if self.user else ""))
        text, n = re.subn(r'#include\s*"hipacc\.hpp"', prologue, text, count=1)
        if n != 1:
            raise EmitError('no #include "hipacc.hpp"')
        header = ("// Emitted by hipacc_amd.emit from " + self.path.split("/")[-1] +
                  " for the MI355X runtime (hipacc_amd/runtime/hipacc_hip.hpp).\n")
        return header + text


def _describe(k: KernelClass) -> str:
    op = k.op
    if k.binning_text:
        extra = "; binning() and reduce() as device code (hipaccBinUser)"
    elif k.global_reduce == "USER":
        extra = "; reduce() as device code (hipaccReduceUser)"
    elif k.global_reduce:
        extra = f"; global {k.global_reduce} reduction (hipaccApplyReduction)"
    else:
        extra = ""
    return _describe_kernel(k) + extra


def _describe_kernel(k: KernelClass) -> str:
    op = k.op
    if isinstance(op, UserOp):
        return "device code, the kernel() body compiled as written (hipaccLaunchUser)"
    if isinstance(op, PointOp):
        return f"point operator {op.op} (hipaccLaunchPoint)"
    if isinstance(op, BilateralOp):
        return (f"bilateral filter over domain {op.domain}, spatial weights {op.mask}"
                f"{', rounded' if op.round else ''} (hipaccLaunchBilateral)")
    return (f"local operator {op.reduce} over {'mask ' + op.mask if op.mask else 'domain ' + op.domain}"
            f"{', rounded' if op.round else ''} (hipaccLaunchLocal)")


def emit(path: str, include_dirs: Sequence[str] = (), defines: Sequence[str] = (),
         fp_contract: bool = True) -> str:
    em = Emitter(path, include_dirs, defines, fp_contract)
    try:
        return em.run()
    finally:
        em.tu.close()
