"""The C-ABI library loads and exports every symbol its headers declare.
No GPU call is made here."""
import ctypes
import os
import re

from gpusharesat_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = ["gss.h", "gss_sharer.h", "gss_residency.h", "gss_ceilings.h", "gss_serve.h"]


def declared_symbols():
    names = set()
    for h in HEADERS:
        path = os.path.join(ROOT, "include", h)
        if not os.path.exists(path):
            continue
        text = open(path).read()
        names |= set(re.findall(r"^\s*(?:int|const char \*|void|int64_t|long)\s*\**\s*"
                                r"(gss_\w+)\s*\(", text, re.M))
    return names


def test_headers_declare_entry_points():
    names = declared_symbols()
    assert "gss_sweep" in names and "gss_engine_create" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in sorted(declared_symbols()) if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_headers():
    assert declared_symbols() == set(_lib.SIGNATURES)


def test_loader_sets_signatures():
    lib = _lib.load()
    assert lib.gss_last_error() == b""


_C_KIND = {"int": "int", "int32_t": "int", "int64_t": "i64", "uint64_t": "u64", "size_t": "u64",
           "float": "float", "double": "double", "uint32_t": "u32", "gss_log_fn": "ptr"}
_CTYPES_KIND = {ctypes.c_int: "int", ctypes.c_int64: "i64", ctypes.c_uint64: "u64",
                ctypes.c_void_p: "ptr", ctypes.c_char_p: "ptr", ctypes.c_float: "float",
                ctypes.c_double: "double", ctypes.c_uint32: "u32"}


def _param_kind(decl: str) -> str:
    decl = decl.strip()
    if "*" in decl or "[" in decl:
        return "ptr"
    words = re.sub(r"\b(const|struct)\b", "", decl).split()
    base = " ".join(words[:-1]) if len(words) > 1 else words[0]
    return _C_KIND.get(base, "unknown " + base)


def test_binding_argument_types_match_headers():
    """Each ctypes argument list has the C parameter kinds of its prototype,
    so that no argument is truncated or widened on the way in."""
    protos = {}
    for h in HEADERS:
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for name, params in re.findall(r"\b(?:int|const char \*|void|double|float)\s*\**\s*"
                                       r"(gss_\w+)\s*\(([^)]*)\)\s*;", text):
            protos[name] = [_param_kind(p) for p in params.split(",")
                            if p.strip() and p.strip() != "void"]
    assert set(protos) == set(_lib.SIGNATURES)
    bad = {n: (want, [_CTYPES_KIND.get(t, "?") for t in _lib.SIGNATURES[n]])
           for n, want in protos.items()
           if [_CTYPES_KIND.get(t, "?") for t in _lib.SIGNATURES[n]] != want}
    assert not bad, bad
