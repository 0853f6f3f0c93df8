"""Minimal ctypes binding of libclang's C API (clang-c/Index.h), enough to
walk a translation unit's cursors, read types and source extents, and
tokenize ranges.  The C API is stable across LLVM versions; cursor kinds are
compared by their spellings (clang_getCursorKindSpelling), not by number.
"""
from __future__ import annotations

import ctypes
import glob
import os
from ctypes import CFUNCTYPE, POINTER, Structure, byref, c_char_p, c_int, c_uint, c_void_p
from typing import Iterator, List, Optional, Sequence, Tuple

LIB_CANDIDATES = ("/opt/rocm/llvm/lib/libclang.so", "libclang.so")


class CXString(Structure):
    _fields_ = [("data", c_void_p), ("flags", c_uint)]


class CXCursor(Structure):
    _fields_ = [("kind", c_int), ("xdata", c_int), ("data", c_void_p * 3)]


class CXType(Structure):
    _fields_ = [("kind", c_int), ("data", c_void_p * 2)]


class CXSourceRange(Structure):
    _fields_ = [("ptr_data", c_void_p * 2), ("begin_int_data", c_uint), ("end_int_data", c_uint)]


class CXSourceLocation(Structure):
    _fields_ = [("ptr_data", c_void_p * 2), ("int_data", c_uint)]


class CXToken(Structure):
    _fields_ = [("int_data", c_uint * 4), ("ptr_data", c_void_p)]


VISITOR = CFUNCTYPE(c_int, CXCursor, CXCursor, c_void_p)
_VISIT_CONTINUE = 1

_SIGS = {
    "clang_createIndex": (c_void_p, [c_int, c_int]),
    "clang_disposeIndex": (None, [c_void_p]),
    "clang_parseTranslationUnit": (c_void_p, [c_void_p, c_char_p, POINTER(c_char_p), c_int,
                                              c_void_p, c_uint, c_uint]),
    "clang_disposeTranslationUnit": (None, [c_void_p]),
    "clang_getTranslationUnitCursor": (CXCursor, [c_void_p]),
    "clang_getNumDiagnostics": (c_uint, [c_void_p]),
    "clang_getDiagnostic": (c_void_p, [c_void_p, c_uint]),
    "clang_getDiagnosticSeverity": (c_int, [c_void_p]),
    "clang_formatDiagnostic": (CXString, [c_void_p, c_uint]),
    "clang_defaultDiagnosticDisplayOptions": (c_uint, []),
    "clang_disposeDiagnostic": (None, [c_void_p]),
    "clang_getCString": (c_char_p, [CXString]),
    "clang_disposeString": (None, [CXString]),
    "clang_getCursorKindSpelling": (CXString, [c_int]),
    "clang_getCursorSpelling": (CXString, [CXCursor]),
    "clang_getCursorType": (CXType, [CXCursor]),
    "clang_getTypeSpelling": (CXString, [CXType]),
    "clang_getCanonicalType": (CXType, [CXType]),
    "clang_getCursorExtent": (CXSourceRange, [CXCursor]),
    "clang_getCursorLocation": (CXSourceLocation, [CXCursor]),
    "clang_getCursorReferenced": (CXCursor, [CXCursor]),
    "clang_Cursor_isNull": (c_int, [CXCursor]),
    "clang_Location_isFromMainFile": (c_int, [CXSourceLocation]),
    "clang_getRangeStart": (CXSourceLocation, [CXSourceRange]),
    "clang_getRangeEnd": (CXSourceLocation, [CXSourceRange]),
    "clang_getSpellingLocation": (None, [CXSourceLocation, POINTER(c_void_p), POINTER(c_uint),
                                         POINTER(c_uint), POINTER(c_uint)]),
    "clang_visitChildren": (c_uint, [CXCursor, VISITOR, c_void_p]),
    "clang_tokenize": (None, [c_void_p, CXSourceRange, POINTER(POINTER(CXToken)), POINTER(c_uint)]),
    "clang_getTokenSpelling": (CXString, [c_void_p, CXToken]),
    "clang_getTokenKind": (c_int, [CXToken]),
    "clang_disposeTokens": (None, [c_void_p, POINTER(CXToken), c_uint]),
}

_LIB = None


def lib() -> ctypes.CDLL:
    global _LIB
    if _LIB is None:
        err = None
        for cand in LIB_CANDIDATES:
            try:
                L = ctypes.CDLL(cand)
                break
            except OSError as e:
                err = e
        else:
            raise OSError(f"libclang not found ({err})")
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype, f.argtypes = res, args
        _LIB = L
    return _LIB


def _s(x: CXString) -> str:
    L = lib()
    r = L.clang_getCString(x)
    L.clang_disposeString(x)
    return r.decode() if r else ""


def resource_dir() -> Optional[str]:
    """The clang resource directory (built-in headers) of the libclang used."""
    for cand in LIB_CANDIDATES:
        base = os.path.dirname(cand)
        dirs = sorted(glob.glob(os.path.join(base, "clang", "*")))
        if dirs:
            return dirs[-1]
    return None


class Cursor:
    """A cursor with the accessors the emitter uses."""

    def __init__(self, tu: "TranslationUnit", c: CXCursor):
        self.tu, self.c = tu, c

    @property
    def kind(self) -> str:
        return _s(lib().clang_getCursorKindSpelling(self.c.kind))

    @property
    def spelling(self) -> str:
        return _s(lib().clang_getCursorSpelling(self.c))

    @property
    def type(self) -> str:
        return _s(lib().clang_getTypeSpelling(lib().clang_getCursorType(self.c)))

    @property
    def canonical_type(self) -> str:
        L = lib()
        return _s(L.clang_getTypeSpelling(L.clang_getCanonicalType(L.clang_getCursorType(self.c))))

    def referenced(self) -> Optional["Cursor"]:
        r = lib().clang_getCursorReferenced(self.c)
        return None if lib().clang_Cursor_isNull(r) else Cursor(self.tu, r)

    def in_main_file(self) -> bool:
        return bool(lib().clang_Location_isFromMainFile(lib().clang_getCursorLocation(self.c)))

    def extent(self) -> Tuple[int, int]:
        """Byte offsets [begin, end) of the cursor's extent in its file."""
        L = lib()
        r = L.clang_getCursorExtent(self.c)
        return _offset(L.clang_getRangeStart(r)), _offset(L.clang_getRangeEnd(r))

    def children(self) -> List["Cursor"]:
        out: List[Cursor] = []

        def visit(c, parent, data):
            out.append(Cursor(self.tu, c))
            return _VISIT_CONTINUE

        lib().clang_visitChildren(self.c, VISITOR(visit), None)
        return out

    def walk(self) -> Iterator["Cursor"]:
        for ch in self.children():
            yield ch
            yield from ch.walk()

    def tokens(self) -> List[str]:
        L = lib()
        toks = POINTER(CXToken)()
        n = c_uint()
        L.clang_tokenize(self.tu.tu, L.clang_getCursorExtent(self.c), byref(toks), byref(n))
        out = [_s(L.clang_getTokenSpelling(self.tu.tu, toks[i])) for i in range(n.value)]
        if n.value:
            L.clang_disposeTokens(self.tu.tu, toks, n)
        return out


def _offset(loc: CXSourceLocation) -> int:
    f = c_void_p()
    line, col, off = c_uint(), c_uint(), c_uint()
    lib().clang_getSpellingLocation(loc, byref(f), byref(line), byref(col), byref(off))
    return off.value


class ParseError(RuntimeError):
    pass


# CXTranslationUnit_DetailedPreprocessingRecord: macro definitions become
# cursors ("macro definition").
DETAILED_PREPROCESSING = 0x01


class TranslationUnit:
    def __init__(self, path: str, args: Sequence[str], options: int = 0):
        L = lib()
        self.index = L.clang_createIndex(0, 0)
        argv = [a.encode() for a in args]
        arr = (c_char_p * len(argv))(*argv)
        self.tu = L.clang_parseTranslationUnit(self.index, path.encode(), arr, len(argv), None, 0,
                                               options)
        if not self.tu:
            raise ParseError(f"libclang could not parse {path}")
        errors = []
        opts = L.clang_defaultDiagnosticDisplayOptions()
        for i in range(L.clang_getNumDiagnostics(self.tu)):
            d = L.clang_getDiagnostic(self.tu, i)
            if L.clang_getDiagnosticSeverity(d) >= 3:      # error, fatal
                errors.append(_s(L.clang_formatDiagnostic(d, opts)))
            L.clang_disposeDiagnostic(d)
        self.errors = errors

    @property
    def cursor(self) -> Cursor:
        return Cursor(self, lib().clang_getTranslationUnitCursor(self.tu))

    def close(self) -> None:
        L = lib()
        if self.tu:
            L.clang_disposeTranslationUnit(self.tu)
            self.tu = None
        if self.index:
            L.clang_disposeIndex(self.index)
            self.index = None
