"""The C-ABI library loads and exports every symbol its headers declare.
No GPU call is made here."""
import ctypes
import os
import re

from gpusharesat_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = ["gss.h", "gss_sharer.h", "gss_residency.h", "gss_ceilings.h"]


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
