"""Python handle on the device check engine (include/gss.h).

This is the thin host-side mirror used by the tests and the benchmark; the
solver-facing interface is the C++ GpuShare::GpuClauseSharer in
include/gpuShareLib/GpuClauseSharer.h (Python mirror: gpusharesat_amd.sharer).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib

HIT_DTYPE = np.dtype([("row", "<u4"), ("len", "<u2"), ("word", "<u2"),
                      ("false_mask", "<u8"), ("unit_mask", "<u8")])
HIT_ID_DTYPE = np.dtype([("row", "<u4"), ("len", "<u2"), ("word", "<u2")])
HIT_CLAUSE_DTYPE = np.dtype([("row", "<u4"), ("len", "<u2"), ("pad", "<u2"), ("words", "<u8")])
VAR_UPDATE_DTYPE = np.dtype([("var", "<u4"), ("word", "<u4"),
                             ("is_def", "<u8"), ("is_true", "<u8")])
assert HIT_DTYPE.itemsize == 24 and VAR_UPDATE_DTYPE.itemsize == 24 and HIT_ID_DTYPE.itemsize == 8

TUNE_WAVES_PER_CU = 1
TUNE_TASK_CLAUSES = 2
TUNE_MAX_HIT_CHUNKS = 3
TUNE_UNROLL = 4
TUNE_FUSE_LEN = 5
TUNE_SCHEDULE = 6
# names of the settings the autotuner chooses, as it reports them
TUNE_KEYS = {"unroll": TUNE_UNROLL, "fuse_len": TUNE_FUSE_LEN,
             "task_clauses": TUNE_TASK_CLAUSES, "waves_per_cu": TUNE_WAVES_PER_CU,
             "schedule": TUNE_SCHEDULE}
MAX_WORDS = 64


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def round_words(words: int) -> int:
    """Table row width used for `words` assignment words (a power of two)."""
    w = 1
    while w < words:
        w *= 2
    if w > MAX_WORDS:
        raise ValueError(f"at most {MAX_WORDS * 64} assignment slots per sweep")
    return w


class Engine:
    def __init__(self, device: int = 0):
        self.device = device
        self._h = ctypes.c_void_p()
        _lib.call("gss_engine_create", device, ctypes.byref(self._h))
        self.words = 1
        self.vars = 0
        self._dependents = []  # weak references to objects built on this engine

    # -- lifetime -------------------------------------------------------------
    def add_dependent(self, obj) -> None:
        """obj.close() runs before the engine is destroyed (whatever order the
        garbage collector finds them in)."""
        import weakref
        self._dependents.append(weakref.ref(obj))

    def close(self) -> None:
        if self._h:
            for ref in self._dependents:
                obj = ref()
                if obj is not None:
                    obj.close()
            self._dependents = []
            _lib.call("gss_engine_destroy", self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def cu_count(self) -> int:
        v = ctypes.c_int()
        _lib.call("gss_engine_cu_count", self._h, ctypes.byref(v))
        return v.value

    def sync(self) -> None:
        _lib.call("gss_engine_sync", self._h)

    def set_page_locked_limit(self, nbytes: int) -> None:
        """Copy-endpoint host buffers of at least nbytes become pageable."""
        _lib.call("gss_engine_set_page_locked_limit", self._h, nbytes)

    # -- clause DB ------------------------------------------------------------
    def append(self, lits: np.ndarray, lens: np.ndarray) -> np.ndarray:
        """Appends clauses (CSR-free: lens[i] literals each, concatenated).
        Returns each clause's row inside its length bucket."""
        lits = np.ascontiguousarray(lits, dtype=np.uint32)
        lens = np.ascontiguousarray(lens, dtype=np.int32)
        if int(lens.sum(dtype=np.int64)) != lits.size:
            raise ValueError("sum(lens) != len(lits)")
        rows = np.empty(lens.size, dtype=np.uint32)
        _lib.call("gss_db_append", self._h, lens.size, _ptr(lits), _ptr(lens), _ptr(rows))
        return rows

    def set_bucket(self, length: int, lits: np.ndarray) -> None:
        lits = np.ascontiguousarray(lits, dtype=np.uint32)
        n = lits.size // length
        _lib.call("gss_db_set_bucket", self._h, length, n, _ptr(lits))

    def bucket_count(self, length: int) -> int:
        v = ctypes.c_int64()
        _lib.call("gss_db_bucket_count", self._h, length, ctypes.byref(v))
        return v.value

    def totals(self) -> tuple[int, int]:
        c, l = ctypes.c_int64(), ctypes.c_int64()
        _lib.call("gss_db_totals", self._h, ctypes.byref(c), ctypes.byref(l))
        return c.value, l.value

    def read_bucket(self, length: int, row0: int = 0, n: int | None = None) -> np.ndarray:
        if n is None:
            n = self.bucket_count(length) - row0
        out = np.empty(n * length, dtype=np.uint32)
        _lib.call("gss_db_read", self._h, length, row0, n, _ptr(out))
        return out.reshape(n, length)

    def clear(self) -> None:
        _lib.call("gss_db_clear", self._h)

    def enable_tags(self) -> None:
        _lib.call("gss_db_enable_tags", self._h)

    def append_tagged(self, lits: np.ndarray, lens: np.ndarray, tags: np.ndarray) -> np.ndarray:
        lits = np.ascontiguousarray(lits, dtype=np.uint32)
        lens = np.ascontiguousarray(lens, dtype=np.int32)
        tags = np.ascontiguousarray(tags, dtype=np.uint32)
        if int(lens.sum(dtype=np.int64)) != lits.size or tags.size != lens.size:
            raise ValueError("sum(lens) != len(lits) or one tag per clause missing")
        rows = np.empty(lens.size, dtype=np.uint32)
        _lib.call("gss_db_append_tagged", self._h, lens.size, _ptr(lits), _ptr(lens), _ptr(tags),
                  _ptr(rows))
        return rows

    def read_tags(self, length: int) -> np.ndarray:
        n = self.bucket_count(length)
        out = np.empty(n, dtype=np.uint32)
        if n:
            _lib.call("gss_db_read_tags", self._h, length, 0, n, _ptr(out))
        return out

    def compact(self, keep: np.ndarray) -> None:
        """Keeps the rows whose tag t has keep[t] set (tags >= keep.size kept)."""
        keep = np.ascontiguousarray(keep, dtype=np.uint8)
        _lib.call("gss_db_compact", self._h, _ptr(keep) if keep.size else None, keep.size)

    # -- assignment table -----------------------------------------------------
    def reset_assignments(self, var_count: int, words: int) -> None:
        w = round_words(words)
        _lib.call("gss_assig_reset", self._h, var_count, w)
        self.words, self.vars = w, var_count

    def grow_vars(self, var_count: int) -> None:
        _lib.call("gss_assig_grow_vars", self._h, var_count)
        self.vars = var_count

    def upload(self, is_def: np.ndarray, is_true: np.ndarray) -> None:
        """Planes are [vars, words] uint64 (words <= table width; missing words
        are undefined)."""
        d = self._pad(is_def)
        t = self._pad(is_true)
        _lib.call("gss_assig_upload", self._h, _ptr(d), _ptr(t))

    def _pad(self, plane: np.ndarray) -> np.ndarray:
        plane = np.asarray(plane, dtype=np.uint64)
        if plane.shape[0] != self.vars:
            raise ValueError("plane has wrong variable count")
        if plane.shape[1] == self.words:
            return np.ascontiguousarray(plane)
        out = np.zeros((self.vars, self.words), dtype=np.uint64)
        out[:, :plane.shape[1]] = plane
        return out

    def download(self) -> tuple[np.ndarray, np.ndarray]:
        d = np.empty((self.vars, self.words), dtype=np.uint64)
        t = np.empty((self.vars, self.words), dtype=np.uint64)
        _lib.call("gss_assig_download", self._h, _ptr(d), _ptr(t))
        return d, t

    def apply(self, ups: np.ndarray) -> None:
        ups = np.ascontiguousarray(ups, dtype=VAR_UPDATE_DTYPE)
        _lib.call("gss_assig_apply", self._h, ups.size, _ptr(ups))

    def roll_forward(self, ups: np.ndarray, last_pos) -> None:
        ups = np.ascontiguousarray(ups, dtype=VAR_UPDATE_DTYPE)
        lp = np.full(self.words, -1, dtype=np.int8)
        lp[:len(last_pos)] = last_pos
        _lib.call("gss_assig_roll_forward", self._h, ups.size, _ptr(ups), _ptr(lp))

    # -- sweep ----------------------------------------------------------------
    def _valid(self, valid) -> np.ndarray:
        v = np.zeros(self.words, dtype=np.uint64)
        valid = np.asarray(valid, dtype=np.uint64).ravel()
        if valid.size > self.words:
            raise ValueError("more valid words than table words")
        v[:valid.size] = valid
        return v

    def sweep_async(self, valid) -> None:
        self._valid_buf = self._valid(valid)
        _lib.call("gss_sweep_async", self._h, _ptr(self._valid_buf))

    def sweep_wait(self) -> int:
        n = ctypes.c_int64()
        _lib.call("gss_sweep_wait", self._h, ctypes.byref(n))
        return n.value

    def sweep(self, valid) -> int:
        v = self._valid(valid)
        n = ctypes.c_int64()
        _lib.call("gss_sweep", self._h, _ptr(v), ctypes.byref(n))
        return n.value

    def hits(self) -> np.ndarray:
        n = ctypes.c_int64()
        _lib.call("gss_hits_copy", self._h, None, 0, ctypes.byref(n))
        out = np.empty(n.value, dtype=HIT_DTYPE)
        if n.value:
            _lib.call("gss_hits_copy", self._h, _ptr(out), n.value, ctypes.byref(n))
        return out

    def kernel_ms(self) -> float:
        v = ctypes.c_float()
        _lib.call("gss_sweep_kernel_ms", self._h, ctypes.byref(v))
        return v.value

    def kernel_stats(self, reset: bool = False) -> tuple[float, int]:
        t, c = ctypes.c_double(), ctypes.c_int64()
        _lib.call("gss_sweep_kernel_stats", self._h, ctypes.byref(t), ctypes.byref(c), int(reset))
        return t.value, c.value

    def set_tuning(self, key: int, value: int) -> None:
        _lib.call("gss_set_tuning", self._h, key, value)

    def apply_tuning(self, settings: dict) -> None:
        """Sets named settings, as autotune() returns them."""
        for name, value in settings.items():
            if name not in TUNE_KEYS:
                raise ValueError(f"unknown tuning setting {name!r}")
            self.set_tuning(TUNE_KEYS[name], int(value))

    def autotune(self, valid, reps: int = 2) -> tuple[dict, float]:
        """Picks the fastest sweep settings for the current DB and table."""
        v = self._valid(valid)
        best = (ctypes.c_int * 5)()
        ms = ctypes.c_float()
        _lib.call("gss_autotune", self._h, _ptr(v), reps, best, ctypes.byref(ms))
        return dict(zip(TUNE_KEYS, list(best))), ms.value

    # -- batched steps --------------------------------------------------------
    def step_enqueue(self, lits=None, lens=None, ups=None, valid=0, last_pos=None,
                     screen: bool = False) -> np.ndarray:
        """Enqueues one step (append, apply, sweep, roll forward) as a graph
        replay; returns the bucket rows of the appended clauses.  screen: the
        step screens clauses against the table's aggregate first
        (GSS_STEP_SCREEN; same records)."""
        if screen:
            from . import serving
            lits = np.ascontiguousarray(np.zeros(0) if lits is None else lits, dtype=np.uint32)
            lens = np.ascontiguousarray(np.zeros(0) if lens is None else lens, dtype=np.int32)
            if int(lens.sum(dtype=np.int64)) != lits.size:
                raise ValueError("sum(lens) != len(lits)")
            rows = np.empty(lens.size, dtype=np.uint32)
            serving.submit(self, self._valid(valid), lits=lits, lens=lens, ups=ups,
                           last_pos=last_pos, flags=serving.STEP_SCREEN, out_rows=rows)
            return rows
        lits = np.ascontiguousarray(np.zeros(0) if lits is None else lits, dtype=np.uint32)
        lens = np.ascontiguousarray(np.zeros(0) if lens is None else lens, dtype=np.int32)
        if int(lens.sum(dtype=np.int64)) != lits.size:
            raise ValueError("sum(lens) != len(lits)")
        ups = np.ascontiguousarray(np.zeros(0, VAR_UPDATE_DTYPE) if ups is None else ups,
                                   dtype=VAR_UPDATE_DTYPE)
        v = self._valid(valid)
        lp = None
        if last_pos is not None:
            lp = np.full(self.words, -1, dtype=np.int8)
            lp[:len(last_pos)] = last_pos
        rows = np.empty(lens.size, dtype=np.uint32)
        _lib.call("gss_step_enqueue", self._h, lens.size, _ptr(lits) if lits.size else None,
                  _ptr(lens) if lens.size else None, _ptr(rows) if rows.size else None,
                  ups.size, _ptr(ups) if ups.size else None, _ptr(v),
                  None if lp is None else _ptr(lp))
        return rows

    def step_wait(self) -> tuple[int, float]:
        n, ms = ctypes.c_int64(), ctypes.c_float()
        _lib.call("gss_step_wait", self._h, ctypes.byref(n), ctypes.byref(ms))
        return n.value, ms.value

    def step_hits_view(self) -> np.ndarray:
        """Records of the last waited step as a view of the engine's pinned
        memory, valid until the next step_wait of that step's slot (two
        waits later; enqueueing that slot again meanwhile is fine)."""
        p, n = ctypes.c_void_p(), ctypes.c_int64()
        _lib.call("gss_step_hits_map", self._h, ctypes.byref(p), ctypes.byref(n))
        if n.value == 0:
            return np.zeros(0, dtype=HIT_DTYPE)
        buf = (ctypes.c_char * (n.value * HIT_DTYPE.itemsize)).from_address(p.value)
        return np.frombuffer(buf, dtype=HIT_DTYPE)

    def step_ids_view(self) -> np.ndarray:
        """gss_hit_id records of the last waited step (submitted with
        GSS_RECORDS_IDS) as a view of pinned memory, valid as step_hits_view's."""
        p, n = ctypes.c_void_p(), ctypes.c_int64()
        _lib.call("gss_step_ids_map", self._h, ctypes.byref(p), ctypes.byref(n))
        if n.value == 0:
            return np.zeros(0, dtype=HIT_ID_DTYPE)
        buf = (ctypes.c_char * (n.value * HIT_ID_DTYPE.itemsize)).from_address(p.value)
        return np.frombuffer(buf, dtype=HIT_ID_DTYPE)

    def step_clauses_view(self) -> np.ndarray:
        """gss_hit_clause records of the last waited step (submitted with
        GSS_RECORDS_CLAUSES) as a view of pinned memory, valid as
        step_hits_view's."""
        p, n = ctypes.c_void_p(), ctypes.c_int64()
        _lib.call("gss_step_clauses_map", self._h, ctypes.byref(p), ctypes.byref(n))
        if n.value == 0:
            return np.zeros(0, dtype=HIT_CLAUSE_DTYPE)
        buf = (ctypes.c_char * (n.value * HIT_CLAUSE_DTYPE.itemsize)).from_address(p.value)
        return np.frombuffer(buf, dtype=HIT_CLAUSE_DTYPE)

    def step_stats(self) -> dict:
        """gss_step_counters of the engine: steps re-run because their records
        overflowed the hit buffer, chunks of 1024 records the step hit
        buffers hold, pinned record-buffer growths in step waits and their
        time."""
        return step_counters(self._h)

    def step_hits(self) -> np.ndarray:
        n = ctypes.c_int64()
        _lib.call("gss_step_hits_copy", self._h, None, 0, ctypes.byref(n))
        out = np.empty(n.value, dtype=HIT_DTYPE)
        if n.value:
            _lib.call("gss_step_hits_copy", self._h, _ptr(out), n.value, ctypes.byref(n))
        return out


class ClauseMap:
    """Maps (bucket length, row) of hit records back to caller clause ids."""

    def __init__(self, lens: np.ndarray, rows: np.ndarray, ids: np.ndarray | None = None):
        lens = np.asarray(lens, dtype=np.int64)
        rows = np.asarray(rows, dtype=np.int64)
        ids = np.arange(lens.size, dtype=np.int64) if ids is None else np.asarray(ids, np.int64)
        maxlen = int(lens.max()) if lens.size else 0
        counts = np.bincount(lens, minlength=maxlen + 2)
        self.base = np.zeros(maxlen + 2, dtype=np.int64)
        self.base[1:] = np.cumsum(counts)[:-1]
        self.table = np.full(lens.size, -1, dtype=np.int64)
        self.table[self.base[lens] + rows] = ids

    def ids(self, hits: np.ndarray) -> np.ndarray:
        return self.table[self.base[hits["len"].astype(np.int64)] + hits["row"].astype(np.int64)]


class _StepCounters(ctypes.Structure):
    _fields_ = [("reruns", ctypes.c_int64), ("hit_chunks", ctypes.c_int64),
                ("host_grows", ctypes.c_int64), ("host_grow_ms", ctypes.c_double),
                ("host_grow_ms_max", ctypes.c_double)]


def step_counters(handle) -> dict:
    """gss_step_stats of an engine handle, as a dict."""
    c = _StepCounters()
    _lib.call("gss_step_stats", handle, ctypes.byref(c))
    return {name: getattr(c, name) for name, _ in _StepCounters._fields_}
