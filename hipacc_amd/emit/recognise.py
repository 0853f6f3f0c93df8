"""Recognition of a DSL kernel's operator from the tokens of its kernel()
and reduce() bodies.

A small expression parser turns the body into a tree; the tree is matched
against the operator forms this library executes (the forms the reference's
operator samples and its code generator's Convolution / Reduce paths use,
lib/AST/Convolution.cpp:228-467, dsl/kernel.hpp:241-296):

  local   output() = [(T)](convolve(M, Reduce::R, [&]() -> A { return M() * in(M); })
                           [+ 0.5f] [/ c])
          output() = [(T)]reduce(D, Reduce::R, [&]() -> A { return in(D); })
          output() = [(T)]reduce(D, Reduce::R, [&]() -> A { return M(D) * in(D); })
  point   output() = in();  output() = in() * a + b (and a * in() + b, in() + b,
          in() - b, in() * a);  output() = p() OP q() with OP in + - *;
          a local variable initialised from the input and updated with
          `v += c` / `v -= c` / `v *= c` before `output() = v`
  global  reduce(l, r) { return min(l, r) | max(l, r) | l + r | l * r; }
  bilateral  the Bilateral_Filter sample's kernel (samples-public/3_Preprocessing/
          Bilateral_Filter/src/main.cpp:63-77), statement for statement, with
          any names, rounded to an integer output or written as a float

Anything else raises Unsupported with the reason: the emitter never guesses
an operator whose arithmetic it cannot reproduce exactly.
"""
from __future__ import annotations

import re
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

TYPE_NAMES = {"uchar", "char", "short", "ushort", "int", "uint", "float", "double", "unsigned",
              "signed", "long"}


class Unsupported(Exception):
    pass


# ---- expressions ---------------------------------------------------------------
# ('num', text) ('id', name) ('call', callee, [args]) ('bin', op, a, b)
# ('neg', a) ('cast', type, a) ('lambda', rettype, expr)

class _Parser:
    def __init__(self, toks: List[str]):
        self.t, self.i = toks, 0

    def peek(self, k: int = 0) -> Optional[str]:
        j = self.i + k
        return self.t[j] if j < len(self.t) else None

    def take(self, want: Optional[str] = None) -> str:
        tok = self.peek()
        if tok is None or (want is not None and tok != want):
            raise Unsupported(f"expected {want!r}, found {tok!r}")
        self.i += 1
        return tok

    def expr(self):
        a = self.mul()
        while self.peek() in ("+", "-"):
            op = self.take()
            a = ("bin", op, a, self.mul())
        return a

    def mul(self):
        a = self.unary()
        while self.peek() in ("*", "/"):
            op = self.take()
            a = ("bin", op, a, self.unary())
        return a

    def _type_at(self, k: int) -> Tuple[Optional[str], int]:
        """A parenthesised type name starting after '(' at offset k."""
        j, words = k, []
        while self.peek(j) in TYPE_NAMES:
            words.append(self.peek(j))
            j += 1
        if words and self.peek(j) == ")":
            return " ".join(words), j + 1 - k
        return None, 0

    def unary(self):
        if self.peek() == "-":
            self.take()
            return ("neg", self.unary())
        if self.peek() == "(":
            ty, n = self._type_at(1)
            if ty:
                self.i += 1 + n
                return ("cast", ty, self.unary())
        return self.postfix()

    def postfix(self):
        a = self.primary()
        while self.peek() == "(":
            self.take("(")
            args = []
            if self.peek() != ")":
                args.append(self.arg())
                while self.peek() == ",":
                    self.take(",")
                    args.append(self.arg())
            self.take(")")
            a = ("call", a, args)
        return a

    def arg(self):
        if self.peek() == "[":
            return self.lam()
        return self.expr()

    def lam(self):
        self.take("[")
        while self.peek() != "]":
            self.take()
        self.take("]")
        self.take("(")
        self.take(")")
        ret = None
        if self.peek() == "->":
            self.take("->")
            words = []
            while self.peek() in TYPE_NAMES:
                words.append(self.take())
            ret = " ".join(words)
        self.take("{")
        self.take("return")
        e = self.expr()
        self.take(";")
        self.take("}")
        return ("lambda", ret, e)

    def primary(self):
        tok = self.peek()
        if tok is None:
            raise Unsupported("unexpected end of expression")
        if tok == "(":
            self.take("(")
            e = self.expr()
            self.take(")")
            return e
        if re.match(r"^[0-9.]", tok):
            return ("num", self.take())
        if re.match(r"^[A-Za-z_]\w*$", tok):
            name = self.take()
            while self.peek() == "::":
                self.take("::")
                name += "::" + self.take()
            return ("id", name)
        raise Unsupported(f"token {tok!r}")


def parse_expr(toks: List[str]):
    p = _Parser(toks)
    e = p.expr()
    if p.i != len(toks):
        raise Unsupported(f"trailing tokens {toks[p.i:]}")
    return e


def _num(e) -> Optional[str]:
    """A numeric literal (possibly negated) as C++ text, else None."""
    if e[0] == "num":
        return e[1]
    if e[0] == "neg" and e[1][0] == "num":
        return "-" + e[1][1]
    return None


def _call_of(e) -> Optional[Tuple[str, List]]:
    if e[0] == "call" and e[1][0] == "id":
        return e[1][1], e[2]
    return None


# ---- operators ---------------------------------------------------------------------
@dataclass
class LocalOp:
    accessor: str
    mask: Optional[str]          # weights (a Mask member), None: unweighted
    domain: Optional[str]        # holes (a Domain member), None: the mask's full extent
    reduce: str                  # SUM MIN MAX PROD
    acc_type: str                # the lambda's type: float -> F32 accumulator, else I32
    round: bool = False
    post_div: Optional[str] = None
    kind: str = "local"


@dataclass
class PointOp:
    op: str                      # COPY AFFINE ADD SUB MUL
    inputs: List[str] = field(default_factory=list)
    alpha: str = "1.0f"
    beta: str = "0.0f"
    kind: str = "point"


@dataclass
class BilateralOp:
    accessor: str
    mask: str                    # spatial weights (a Mask member)
    domain: str                  # the positions iterated (a Domain member)
    sigma_r: str                 # the scalar member holding sigma_r
    round: bool                  # (T)(p / d + 0.5f) instead of p / d
    kind: str = "bilateral"


@dataclass
class GlobalOp:
    reduce: str                  # SUM MIN MAX PROD
    accessor: str = ""
    kind: str = "global"


# The Bilateral_Filter sample's kernel() as tokens; $NAME binds one
# identifier token (the same token wherever the name recurs), $T the output
# type.  {CAST} is "( float )" or nothing, {OUT} the rounded or the float
# output statement.
_BILATERAL = (
    "{ float $c = 0.5f / ( $S * $S ) ; float $d = 0.0f ; float $p = 0.0f ; "
    "float $ctr = $IN ( ) ; "
    "iterate ( $DOM , [ & ] ( ) -> void { "
    "float $diff = {CAST} ( $IN ( $DOM ) ) - $ctr ; "
    "float $s = expf ( - $c * $diff * $diff ) * $MASK ( $DOM ) ; "
    "$d += $s ; $p += $s * $IN ( $DOM ) ; } ) ; "
    "{OUT} }")
_BILATERAL_OUT = {True: "output ( ) = ( $T ) ( $p / $d + 0.5f ) ;",
                  False: "output ( ) = $p / $d ;"}


def _match(template: List[str], toks: List[str]) -> Optional[Dict[str, str]]:
    if len(template) != len(toks):
        return None
    env: Dict[str, str] = {}
    for t, k in zip(template, toks):
        if t.startswith("$"):
            if not re.match(r"^[A-Za-z_]\w*$", k) or env.setdefault(t, k) != k:
                return None
        elif t != k:
            return None
    return env


def classify_bilateral(body: List[str], accessors: List[str], masks: List[str],
                       domains: List[str], scalars: List[str]) -> BilateralOp:
    """The bilateral filter, if body is the Bilateral_Filter sample's kernel."""
    for rnd in (True, False):
        for cast in ("( float )", ""):
            tpl = _BILATERAL.replace("{CAST}", cast).replace("{OUT}", _BILATERAL_OUT[rnd]).split()
            env = _match(tpl, body)
            if env is None:
                continue
            names = [env[k] for k in ("$c", "$d", "$p", "$ctr", "$diff", "$s")]
            if (env["$IN"] in accessors and env["$DOM"] in domains and env["$MASK"] in masks
                    and env["$S"] in scalars and len(set(names)) == len(names)
                    and not set(names) & {env["$IN"], env["$DOM"], env["$MASK"], env["$S"]}
                    and env.get("$T", "uchar") in ("uchar", "unsigned char")):
                return BilateralOp(env["$IN"], env["$MASK"], env["$DOM"], env["$S"], rnd)
    raise Unsupported("not the bilateral filter's form")


def _split_statements(toks: List[str]) -> List[List[str]]:
    if not toks or toks[0] != "{" or toks[-1] != "}":
        raise Unsupported("body is not a compound statement")
    out, cur, depth = [], [], 0
    for tok in toks[1:-1]:
        cur.append(tok)
        if tok in "([{":
            depth += 1
        elif tok in ")]}":
            depth -= 1
        elif tok == ";" and depth == 0:
            out.append(cur[:-1])
            cur = []
    if cur:
        raise Unsupported("statement without ';'")
    return out


def output_expr(body: List[str]):
    """The expression assigned to output() by a kernel() body, with local
    variables substituted (`T v = e;`, `v = e;`, `v op= e;`)."""
    env: Dict[str, tuple] = {}
    result = None
    for st in _split_statements(body):
        if st[:4] == ["output", "(", ")", "="]:
            result = _subst(parse_expr(st[4:]), env)
            continue
        j = 0
        while j < len(st) and st[j] in TYPE_NAMES:
            j += 1
        if j > 0 and len(st) > j + 1 and st[j + 1] == "=":
            env[st[j]] = _subst(parse_expr(st[j + 2:]), env)
        elif len(st) > 2 and st[1] == "=" and st[0] in env:
            env[st[0]] = _subst(parse_expr(st[2:]), env)
        elif len(st) > 2 and st[1] in ("+=", "-=", "*=") and st[0] in env:
            env[st[0]] = ("bin", st[1][0], env[st[0]], _subst(parse_expr(st[2:]), env))
        else:
            raise Unsupported(f"statement {' '.join(st)!r}")
    if result is None:
        raise Unsupported("no output() assignment")
    return result


def _subst(e, env):
    if e[0] == "id" and e[1] in env:
        return env[e[1]]
    if e[0] == "bin":
        return ("bin", e[1], _subst(e[2], env), _subst(e[3], env))
    if e[0] in ("neg",):
        return ("neg", _subst(e[1], env))
    if e[0] == "cast":
        return ("cast", e[1], _subst(e[2], env))
    if e[0] == "call":
        return ("call", e[1], [_subst(a, env) for a in e[2]])
    return e


def classify_kernel(body: List[str], accessors: List[str], masks: List[str],
                    domains: List[str]):
    """The operator of a kernel() body; accessors / masks / domains are the
    kernel class's members of those types."""
    e = output_expr(body)
    if e[0] == "cast":
        e = e[2]
    rnd, post_div = False, None
    if e[0] == "bin" and e[1] == "+" and _num(e[3]) in ("0.5f", "0.5"):
        rnd, e = True, e[2]
    if e[0] == "bin" and e[1] == "/" and _num(e[3]) is not None:
        post_div, e = _num(e[3]), e[2]
    c = _call_of(e)
    if c and c[0] in ("convolve", "reduce") and len(c[1]) == 3:
        return _local(c[0], c[1], rnd, post_div, accessors, masks, domains)
    if rnd or post_div:
        raise Unsupported("rounding or division outside a local operator")
    return _point(e, accessors)


def _local(fn, args, rnd, post_div, accessors, masks, domains) -> LocalOp:
    dom_or_mask, red, lam = args
    if dom_or_mask[0] != "id" or red[0] != "id" or lam[0] != "lambda":
        raise Unsupported(f"{fn} arguments")
    region = dom_or_mask[1]
    m = re.match(r"^(?:hipacc::)?Reduce::(SUM|MIN|MAX|PROD)$", red[1])
    if not m:
        raise Unsupported(f"reduction {red[1]}")
    acc_type = lam[1] or "float"
    body = lam[2]
    if body[0] == "cast":
        body = body[2]

    def acc_call(x, arg):
        c = _call_of(x)
        return c and c[0] in accessors and len(c[1]) == 1 and c[1][0] == ("id", arg)

    def mask_call(x, arg):
        c = _call_of(x)
        if not c or c[0] not in masks:
            return False
        return (arg is None and not c[1]) or (len(c[1]) == 1 and c[1][0] == ("id", arg))

    if fn == "convolve":
        if region not in masks:
            raise Unsupported("convolve over a non-mask")
        if body[0] == "bin" and body[1] == "*":
            a, b = body[2], body[3]
            for w, x in ((a, b), (b, a)):
                if mask_call(w, None) and acc_call(x, region):
                    return LocalOp(_call_of(x)[0], _call_of(w)[0], None, m.group(1), acc_type,
                                   rnd, post_div)
        raise Unsupported("convolve lambda is not mask() * accessor(mask)")
    # reduce over a Domain (or a Mask used as one)
    if acc_call(body, region):
        return LocalOp(_call_of(body)[0], None, region, m.group(1), acc_type, rnd, post_div)
    if body[0] == "bin" and body[1] == "*":
        a, b = body[2], body[3]
        for w, x in ((a, b), (b, a)):
            if mask_call(w, region) and acc_call(x, region):
                return LocalOp(_call_of(x)[0], _call_of(w)[0], region, m.group(1), acc_type, rnd,
                               post_div)
    raise Unsupported("reduce lambda is neither accessor(dom) nor mask(dom) * accessor(dom)")


def _point(e, accessors) -> PointOp:
    def acc0(x):
        c = _call_of(x)
        return c[0] if c and c[0] in accessors and not c[1] else None

    if e[0] == "cast":
        e = e[2]
    if acc0(e):
        return PointOp("COPY", [acc0(e)])
    if e[0] == "bin":
        op, a, b = e[1], e[2], e[3]
        if acc0(a) and acc0(b) and op in "+-*":
            return PointOp({"+": "ADD", "-": "SUB", "*": "MUL"}[op], [acc0(a), acc0(b)])
        # alpha * x + beta in exactly that evaluation order
        if op in "+-" and _num(b) is not None:
            beta = _num(b) if op == "+" else _neg(_num(b))
            inner = a
            if acc0(inner):
                return PointOp("AFFINE", [acc0(inner)], "1.0f", beta)
            if inner[0] == "bin" and inner[1] == "*":
                for x, k in ((inner[2], inner[3]), (inner[3], inner[2])):
                    if acc0(x) and _num(k) is not None:
                        return PointOp("AFFINE", [acc0(x)], _num(k), beta)
        if op == "+" and _num(a) is not None and b[0] == "bin" and b[1] == "*":
            for x, k in ((b[2], b[3]), (b[3], b[2])):
                if acc0(x) and _num(k) is not None:
                    return PointOp("AFFINE", [acc0(x)], _num(k), _num(a))
        if op == "*":
            for x, k in ((a, b), (b, a)):
                if acc0(x) and _num(k) is not None:
                    return PointOp("AFFINE", [acc0(x)], _num(k), "0.0f")
    raise Unsupported("point expression is not a copy, alpha * x + beta or a two-image + - *")


def _neg(text: str) -> str:
    return text[1:] if text.startswith("-") else "-" + text


def classify_reduce(body: List[str], params: Tuple[str, str]) -> str:
    """SUM / MIN / MAX / PROD of a Kernel::reduce(left, right) body."""
    sts = _split_statements(body)
    if len(sts) != 1 or not sts[0] or sts[0][0] != "return":
        raise Unsupported("reduce() is not a single return")
    e = parse_expr(sts[0][1:])
    left, right = ("id", params[0]), ("id", params[1])
    c = _call_of(e)
    if c and len(c[1]) == 2 and set(map(repr, c[1])) == {repr(left), repr(right)}:
        name = c[0].split("::")[-1]
        if name in ("min", "fmin", "fminf"):
            return "MIN"
        if name in ("max", "fmax", "fmaxf"):
            return "MAX"
    if e[0] == "bin" and {repr(e[2]), repr(e[3])} == {repr(left), repr(right)}:
        if e[1] == "+":
            return "SUM"
        if e[1] == "*":
            return "PROD"
    # the ternary forms are left to the parser's error
    raise Unsupported("reduce() is not min, max, + or * of its two arguments")
