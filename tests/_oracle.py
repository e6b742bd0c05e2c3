"""ctypes binding of the CPU oracle (oracle/oracle.c) for the tests.

The oracle is test infrastructure: it is the checker, never the thing tested.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "liboracle.so")

HIT_DTYPE = np.dtype([("clause", "<u4"), ("word", "<u4"),
                      ("false_mask", "<u8"), ("unit_mask", "<u8")])

_lib = None


def load() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        src = os.path.join(ORACLE_DIR, "oracle.c")
        if (not os.path.exists(ORACLE_LIB)
                or os.path.getmtime(ORACLE_LIB) < os.path.getmtime(src)):
            subprocess.run(["make", "-C", ORACLE_DIR, "-s"], check=True)
        lib = ctypes.CDLL(ORACLE_LIB)
        P, I, I64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64
        lib.orc_rule_clause.argtypes = [P, I, P, P, I, I, ctypes.c_uint64, P, P]
        lib.orc_rule_clause.restype = None
        lib.orc_sweep.argtypes = [I64, P, P, I, P, P, P, I, P, I64]
        lib.orc_sweep.restype = I64
        lib.orc_perftest_counts.argtypes = [I, I, I, I, I, I, I, P]
        lib.orc_perftest_counts.restype = I
        lib.orc_perftest_clauses.argtypes = [I, I, I, I, ctypes.c_double, I, P, P]
        lib.orc_perftest_clauses.restype = None
        lib.orc_perftest_assigs.argtypes = [I, I, ctypes.c_double, P, P]
        lib.orc_perftest_assigs.restype = None
        lib.orc_drand.argtypes = [P]
        lib.orc_drand.restype = ctypes.c_double
        _lib = lib
    return _lib


def threads() -> int:
    return max(1, min(16, os.cpu_count() or 1))


def sweep(offsets, lits, is_def, is_true, valid, nthreads: int | None = None) -> np.ndarray:
    """All (clause, word) records with a report, sorted by (clause, word)."""
    lib = load()
    offsets = np.ascontiguousarray(offsets, dtype=np.int64)
    lits = np.ascontiguousarray(lits, dtype=np.uint32)
    is_def = np.ascontiguousarray(is_def, dtype=np.uint64)
    is_true = np.ascontiguousarray(is_true, dtype=np.uint64)
    W = is_def.shape[1]
    v = np.zeros(W, dtype=np.uint64)
    valid = np.asarray(valid, dtype=np.uint64).ravel()
    v[:min(W, valid.size)] = valid[:W]
    nt = nthreads or threads()
    n = lib.orc_sweep(offsets.size - 1, offsets.ctypes.data, lits.ctypes.data, W,
                      is_def.ctypes.data, is_true.ctypes.data, v.ctypes.data, nt, None, 0)
    if n < 0:
        raise MemoryError("oracle sweep failed")
    out = np.empty(n, dtype=HIT_DTYPE)
    if n:
        m = lib.orc_sweep(offsets.size - 1, offsets.ctypes.data, lits.ctypes.data, W,
                          is_def.ctypes.data, is_true.ctypes.data, v.ctypes.data, nt,
                          out.ctypes.data, n)
        assert m == n
    return out


def rule(lits, is_def, is_true, w: int, start: int) -> tuple[int, int]:
    """Reference rule on one clause and one word."""
    lib = load()
    lits = np.ascontiguousarray(lits, dtype=np.uint32)
    is_def = np.ascontiguousarray(is_def, dtype=np.uint64)
    is_true = np.ascontiguousarray(is_true, dtype=np.uint64)
    fm, um = ctypes.c_uint64(), ctypes.c_uint64()
    lib.orc_rule_clause(lits.ctypes.data, lits.size, is_def.ctypes.data, is_true.ctypes.data,
                        is_def.shape[1], w, ctypes.c_uint64(start), ctypes.byref(fm),
                        ctypes.byref(um))
    return fm.value, um.value


def perftest_counts(nclauses: int, rounds: int, sign_first: int = 1, min_size: int = 12,
                    max_size: int = 20, nvars: int = 500) -> np.ndarray:
    lib = load()
    c = np.zeros(rounds, dtype=np.int64)
    rc = lib.orc_perftest_counts(nclauses, min_size, max_size, nvars, rounds, sign_first,
                                 threads(), c.ctypes.data)
    if rc:
        raise MemoryError("perftest oracle failed")
    return c


def perftest_instance(nclauses: int, rounds: int, sign_first: int = 1, min_size: int = 12,
                      max_size: int = 20, nvars: int = 500):
    """The perfTest clause set and assignments (as planes), for GPU parity."""
    lib = load()
    off = np.zeros(nclauses + 1, dtype=np.int64)
    lits = np.zeros(nclauses * (max_size - 1), dtype=np.uint32)
    lib.orc_perftest_clauses(nclauses, min_size, max_size, nvars, 0.4, sign_first,
                             off.ctypes.data, lits.ctypes.data)
    lits = lits[:off[-1]].copy()
    W = (rounds + 63) // 64
    d = np.zeros((nvars, W), dtype=np.uint64)
    t = np.zeros((nvars, W), dtype=np.uint64)
    lib.orc_perftest_assigs(rounds, nvars, 0.6, d.ctypes.data, t.ctypes.data)
    return off, lits, d, t
