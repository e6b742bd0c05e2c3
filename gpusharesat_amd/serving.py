"""Serving: requests (sets of assignments from a solver) arrive continuously
and are checked against a device-resident clause DB that grows under an HBM
budget.  The loop itself is native (include/gss_serve.h, GpuServer below):
clients publish requests into a pinned ring, the library's serving thread
keeps batched steps on the device, and each step admits the published
requests that fit on the device when it starts.  This module drives it with
the serving workload and measures it.

A request of k assignments from solver s takes ceil(k / 64) consecutive
table words.  The step loads those words from solver s's base (device-side
packed planes), after the base drift published before it, and then applies
the request's own per-slot differences as masked patches (gss_var_patch).
So the host sends only differences, never whole assignments, and no request
depends on which words it lands in: its records are the same alone or in any
mix.

Workload (serving-32): 32 solvers over 2M variables, each assigning a
fraction U[0.3, 0.7] of them; requests arrive as a Poisson process at 2000
per second, each of U[64, 1024] assignments that differ from the solver's
base in `flips` variables per slot; clauses arrive at 200k per second with
lengths 2 + Geometric(0.1) capped at 200, distinct variables, and are
admitted while the DB fits a 120 GB budget.  Assignments and clauses are
drawn around a planted total assignment sigma (an assigned variable takes
sigma's value with probability `agree`; a clause literal is true under sigma
with probability `lit_true`): uniformly random clauses against uniformly
random assignments would make about a quarter of all (clause, 64-slot word)
pairs a record, hundreds of MB per step, which no solver population sends.
"""
from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .engine import Engine, HIT_DTYPE

PATCH_DTYPE = np.dtype([("var", "<u4"), ("word", "<u4"), ("mask", "<u8"),
                        ("is_def", "<u8"), ("is_true", "<u8")])
BASE_UP_DTYPE = np.dtype([("var", "<u4"), ("base", "<u2"), ("is_def", "u1"), ("is_true", "u1")])
LOAD_DTYPE = np.dtype([("word", "<u2"), ("base", "<u2")])
assert PATCH_DTYPE.itemsize == 32 and BASE_UP_DTYPE.itemsize == 8 and LOAD_DTYPE.itemsize == 4

WORDS = 64
ONES = np.uint64(0xFFFFFFFFFFFFFFFF)


class _StepDesc(ctypes.Structure):
    _fields_ = [("n_clauses", ctypes.c_int64), ("lits", ctypes.c_void_p),
                ("lens", ctypes.c_void_p), ("out_rows", ctypes.c_void_p),
                ("n_ups", ctypes.c_int64), ("ups", ctypes.c_void_p),
                ("n_patches", ctypes.c_int64), ("patches", ctypes.c_void_p),
                ("n_base_ups", ctypes.c_int64), ("base_ups", ctypes.c_void_p),
                ("n_loads", ctypes.c_int64), ("loads", ctypes.c_void_p),
                ("valid", ctypes.c_void_p), ("last_pos", ctypes.c_void_p),
                ("records", ctypes.c_int32), ("flags", ctypes.c_int32)]


def _p(a):
    return a.ctypes.data if a is not None and a.size else None


RECORDS_FULL = 0
RECORDS_IDS = 1
RECORDS_CLAUSES = 2
STEP_SCREEN = 1  # gss_step_desc.flags


def submit(e: Engine, valid, lits=None, lens=None, patches=None, base_ups=None, loads=None,
           ups=None, last_pos=None, records: int = RECORDS_FULL, flags: int = 0,
           out_rows: np.ndarray | None = None) -> None:
    """gss_step_submit with numpy arrays (kept alive by the caller until the
    call returns; the engine copies them into its pinned staging)."""
    d = _StepDesc()
    d.records = records
    d.flags = flags
    if out_rows is not None and out_rows.size:
        d.out_rows = out_rows.ctypes.data
    v = np.zeros(e.words, np.uint64)
    v[:len(valid)] = valid
    arrs = {"valid": v}
    if lens is not None and len(lens):
        arrs["lits"] = np.ascontiguousarray(lits, np.uint32)
        arrs["lens"] = np.ascontiguousarray(lens, np.int32)
        d.n_clauses, d.lits, d.lens = arrs["lens"].size, _p(arrs["lits"]), _p(arrs["lens"])
    given = {"patches": patches, "base_ups": base_ups, "loads": loads}
    for name, dt, cnt, ptr in (("patches", PATCH_DTYPE, "n_patches", "patches"),
                               ("base_ups", BASE_UP_DTYPE, "n_base_ups", "base_ups"),
                               ("loads", LOAD_DTYPE, "n_loads", "loads")):
        a = given[name]
        if a is not None and len(a):
            arrs[name] = np.ascontiguousarray(a, dt)
            setattr(d, cnt, arrs[name].size)
            setattr(d, ptr, _p(arrs[name]))
    if ups is not None and len(ups):
        from .engine import VAR_UPDATE_DTYPE
        arrs["ups"] = np.ascontiguousarray(ups, VAR_UPDATE_DTYPE)
        d.n_ups, d.ups = arrs["ups"].size, _p(arrs["ups"])
    if last_pos is not None:
        lp = np.full(e.words, -1, np.int8)
        lp[:len(last_pos)] = last_pos
        arrs["last_pos"] = lp
        d.last_pos = _p(lp)
    d.valid = _p(v)
    _lib.call("gss_step_submit", e.handle, ctypes.byref(d))


def bases_reset(e: Engine, n: int) -> None:
    _lib.call("gss_bases_reset", e.handle, n)


def slot_patches(rng, var_count: int, k: int, flips: int, p_assigned: float,
                 sigma: np.ndarray | None = None, agree: float = 0.5) -> np.ndarray:
    """Patches (relative words 0 .. ceil(k/64)-1) of k assignments that each
    redraw `flips` variables of the base: assigned with probability
    p_assigned, with the planted value sigma[v] with probability `agree`
    (polarity Bernoulli(0.5) without sigma)."""
    slots = np.repeat(np.arange(k, dtype=np.int64), flips)
    vs = rng.integers(0, var_count, size=slots.size, dtype=np.int64)
    nd = rng.random(slots.size) < p_assigned
    if sigma is None:
        nt = nd & (rng.random(slots.size) < 0.5)
    else:
        nt = nd & ((rng.random(slots.size) < agree) == sigma[vs])
    key = (vs * 16 + slots // 64) * 64 + slots % 64
    uk, ridx = np.unique(key[::-1], return_index=True)  # a repeated draw keeps the last
    sel = slots.size - 1 - ridx
    bit = np.uint64(1) << (uk % 64).astype(np.uint64)
    vw = uk // 64
    g, start = np.unique(vw, return_index=True)
    out = np.zeros(g.size, PATCH_DTYPE)
    out["var"] = g // 16
    out["word"] = g % 16
    out["mask"] = np.bitwise_or.reduceat(bit, start)
    out["is_def"] = np.bitwise_or.reduceat(np.where(nd[sel], bit, np.uint64(0)), start)
    out["is_true"] = np.bitwise_or.reduceat(np.where(nt[sel], bit, np.uint64(0)), start)
    return out


@dataclass
class ServeConfig:
    var_count: int = 2_000_000
    solvers: int = 32
    rate: float = 2000.0            # requests per second (Poisson)
    set_min: int = 64
    set_max: int = 1024
    p_min: float = 0.3
    p_max: float = 0.7
    clause_rate: float = 200_000.0  # clauses per second
    duration: float = 60.0          # seconds of arrivals
    hbm_budget: float = 120e9       # bytes of clause literals admitted
    drift: int = 256                # base variables redrawn between two requests of a solver
    flips: int = 32                 # variables redrawn per assignment
    agree: float = 0.99             # an assigned variable takes the planted value
    lit_true: float = 0.9           # a clause literal is true under the planted assignment
    planted: bool = True            # False: uniform polarities and random clause signs
    pool: int = 48                  # distinct request payloads
    clause_pool: int = 2_000_000
    max_inflight: int = 2           # steps on the device at once
    records: int = RECORDS_FULL     # RECORDS_IDS: clause and word only
    seed: int = 5


@dataclass
class Request:
    idx: int
    t_arrive: float
    solver: int
    k: int
    payload: int


@dataclass
class ServeResult:
    requests: int
    steps: int
    elapsed: float
    checks: float
    latency_ms: np.ndarray
    clauses_end: int
    literals_end: int
    records: int
    max_queue: int
    device_ms: float = 0.0
    arrive_s: np.ndarray = None     # arrival time of the request of each latency
    kept: dict = field(default_factory=dict)  # request id -> (result, records), sampled
    # (seconds, steps, device ms, clauses) of the server, about every 5 s
    samples: list = field(default_factory=list)
    host_grouped: int = 0    # steps whose records the serving thread grouped itself
    graph_captures: int = 0  # step graphs the engine captured
    thread: dict = field(default_factory=dict)  # serving thread's submit / wait host times
    lag_ms: np.ndarray = None  # arrival -> published in the ring (this driver's part)
    step_stats: dict = field(default_factory=dict)  # hit-buffer re-runs and size

    @property
    def checks_per_s(self) -> float:
        return self.checks / self.elapsed


SERVE_HIT_DTYPE = np.dtype([("clause", "<i8"), ("word", "<u4"), ("pad", "<u4"),
                            ("false_mask", "<u8"), ("unit_mask", "<u8")])


class _ServerConfig(ctypes.Structure):
    _fields_ = [("var_count", ctypes.c_int64), ("solvers", ctypes.c_int32),
                ("words", ctypes.c_int32), ("hbm_budget", ctypes.c_double),
                ("max_inflight", ctypes.c_int32), ("records", ctypes.c_int32),
                ("request_ring", ctypes.c_int32), ("pad", ctypes.c_int32),
                ("patch_ring", ctypes.c_int64)]


class _Result(ctypes.Structure):
    _fields_ = [("request", ctypes.c_int64), ("step", ctypes.c_int64),
                ("clauses", ctypes.c_int64), ("base_updates", ctypes.c_int64),
                ("solver", ctypes.c_int32), ("k", ctypes.c_int32),
                ("first_word", ctypes.c_int32), ("pad", ctypes.c_int32),
                ("latency_ms", ctypes.c_double), ("n_records", ctypes.c_int64),
                ("records", ctypes.c_void_p)]


class _Stats(ctypes.Structure):
    _fields_ = [("steps", ctypes.c_int64), ("requests_done", ctypes.c_int64),
                ("requests_pending", ctypes.c_int64), ("records", ctypes.c_int64),
                ("clauses", ctypes.c_int64), ("clauses_refused", ctypes.c_int64),
                ("literals", ctypes.c_int64), ("checks", ctypes.c_double),
                ("device_ms", ctypes.c_double), ("host_grouped", ctypes.c_int64),
                ("graph_captures", ctypes.c_int64), ("submit_ms_total", ctypes.c_double),
                ("submit_ms_max", ctypes.c_double), ("wait_ms_max", ctypes.c_double)]


class GpuServer:
    """ctypes handle on the native serving loop (include/gss_serve.h)."""

    def __init__(self, var_count: int, solvers: int, words: int = WORDS,
                 hbm_budget: float = 120e9, max_inflight: int = 2, records: int = RECORDS_FULL,
                 request_ring: int = 4096, patch_ring: int = 8 << 20, device: int = 0):
        c = _ServerConfig()
        _lib.call("gss_server_default_config", ctypes.byref(c))
        c.var_count, c.solvers, c.words = var_count, solvers, words
        c.hbm_budget, c.max_inflight, c.records = hbm_budget, max_inflight, records
        c.request_ring, c.patch_ring = request_ring, patch_ring
        self.words = words
        self._h = ctypes.c_void_p()
        _lib.call("gss_server_create", device, ctypes.byref(c), ctypes.byref(self._h))

    def close(self) -> None:
        if self._h:
            _lib.call("gss_server_destroy", self._h)
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

    def engine_handle(self):
        e = ctypes.c_void_p()
        _lib.call("gss_server_engine", self._h, ctypes.byref(e))
        return e

    def set_tuning(self, settings: dict) -> None:
        from .engine import TUNE_KEYS
        e = self.engine_handle()
        for name, value in settings.items():
            _lib.call("gss_set_tuning", e, TUNE_KEYS[name], int(value))

    def step_stats(self) -> dict:
        """The engine's step counters (gss_step_stats): re-runs, hit-buffer
        chunks, pinned record-buffer growths and their time."""
        from .engine import step_counters
        return step_counters(self.engine_handle())

    def update_base(self, ups: np.ndarray) -> None:
        ups = np.ascontiguousarray(ups, BASE_UP_DTYPE)
        if ups.size:
            _lib.call("gss_server_update_base", self._h, ups.size, _p(ups))

    def add_clauses(self, lits: np.ndarray, lens: np.ndarray) -> int:
        """First id of the admitted clauses, or -1 if the budget refused them."""
        lits = np.ascontiguousarray(lits, np.uint32)
        lens = np.ascontiguousarray(lens, np.int32)
        first = ctypes.c_int64()
        _lib.call("gss_server_add_clauses", self._h, lens.size, _p(lits), _p(lens),
                  ctypes.byref(first))
        return first.value

    def submit(self, solver: int, k: int, patches: np.ndarray) -> int | None:
        """Request id, or None if a ring is full."""
        patches = np.ascontiguousarray(patches, PATCH_DTYPE)
        rid = ctypes.c_int64()
        rc = _lib.load().gss_server_submit(self._h, solver, k, patches.size, _p(patches),
                                            ctypes.byref(rid))
        if rc == -4 and b"full" in _lib.load().gss_last_error():
            return None
        _lib.check("gss_server_submit", rc)
        return rid.value

    def poll(self, max_ids: int = 1024, timeout_us: int = 0) -> list[int]:
        ids = (ctypes.c_int64 * max_ids)()
        n = ctypes.c_int64()
        _lib.call("gss_server_poll", self._h, max_ids, ids, ctypes.byref(n), int(timeout_us))
        return list(ids[:n.value])

    def result(self, rid: int, records: bool = True) -> tuple[dict, np.ndarray | None]:
        r = _Result()
        _lib.call("gss_server_result", self._h, rid, ctypes.byref(r))
        meta = {f: getattr(r, f) for f, _ in _Result._fields_ if f not in ("records", "pad")}
        recs = None
        if records:
            p, n = ctypes.c_void_p(), ctypes.c_int64()
            _lib.call("gss_server_records", self._h, rid, ctypes.byref(p), ctypes.byref(n))
            if n.value:
                buf = (ctypes.c_char * (n.value * SERVE_HIT_DTYPE.itemsize)).from_address(p.value)
                recs = np.frombuffer(buf, SERVE_HIT_DTYPE).copy()
            else:
                recs = np.zeros(0, SERVE_HIT_DTYPE)
        return meta, recs

    def release(self, rid: int) -> None:
        _lib.call("gss_server_release", self._h, rid)

    def stats(self) -> dict:
        st = _Stats()
        _lib.call("gss_server_stats_get", self._h, ctypes.byref(st))
        return {f: getattr(st, f) for f, _ in _Stats._fields_}


class Server:
    """The serving-32 workload driven through the native loop: arrivals on a
    Poisson schedule (base drift, clause growth and the request published at
    each arrival), completions polled between arrivals."""

    def __init__(self, cfg: ServeConfig, device: int = 0, log=None, keep_history: bool = False):
        from . import workloads
        self.cfg = cfg
        self.log = log or (lambda m: None)
        rng = np.random.default_rng(cfg.seed)
        self.rng = rng
        self.srv = GpuServer(cfg.var_count, cfg.solvers, WORDS, cfg.hbm_budget,
                             cfg.max_inflight, cfg.records, device=device)
        self.device = device
        self.p_assigned = rng.uniform(cfg.p_min, cfg.p_max, size=cfg.solvers)
        self.sigma = rng.random(cfg.var_count) < 0.5  # the planted assignment
        # payload pools: per solver-independent sets of 1024 slots, patches in
        # relative words sorted by word, with the first patch of every word; a
        # request of k assignments sends the patches of its first ceil(k/64)
        # words (slots past k in its last word are not checked)
        agree = cfg.agree if cfg.planted else 0.5
        sigma = self.sigma if cfg.planted else None
        self.pool, self.pool_off = [], []
        for _ in range(cfg.pool):
            pt = slot_patches(rng, cfg.var_count, 1024, cfg.flips, 0.5, sigma, agree)
            pt = pt[np.argsort(pt["word"], kind="stable")]
            self.pool.append(pt)
            self.pool_off.append(np.searchsorted(pt["word"], np.arange(17)))
        self.log(f"payload pool: {sum(p.size for p in self.pool)} patches")
        lens = workloads.clause_lengths(rng, cfg.clause_pool, 0.1, 200)
        self.cl_lens = lens
        self.cl_off = np.zeros(lens.size + 1, np.int64)
        np.cumsum(lens, out=self.cl_off[1:])
        v = workloads.distinct_vars(rng, lens, cfg.var_count)
        if cfg.planted:
            st = rng.random(v.size) < cfg.lit_true
            # literal 2v + 1 is the negation: true under sigma iff sigma(v) is false
            self.cl_lits = ((v << 1) | (st != self.sigma[v])).astype(np.uint32)
        else:
            self.cl_lits = ((v << 1) | rng.integers(0, 2, size=v.size)).astype(np.uint32)
        self.cl_next = 0
        self.keep_history = keep_history
        self.base_history: list[np.ndarray] = []   # base updates in submission order
        self.clause_ids: list[tuple[int, int, int]] = []  # (first id, pool lo, pool hi)

    @property
    def e(self):
        return self.srv

    def tune(self, clauses: int = 2_000_000) -> dict:
        """Chooses the sweep tiles on a scratch engine holding `clauses` pooled
        clauses and a 64-word table, and applies them to the server's engine
        (whose DB starts empty)."""
        from .engine import Engine
        c = self.cfg
        t = Engine(self.device)
        try:
            n = min(clauses, c.clause_pool)
            t.append(self.cl_lits[:self.cl_off[n]], self.cl_lens[:n])
            t.reset_assignments(c.var_count, WORDS)
            bases_reset(t, 1)
            u = np.zeros(c.var_count, BASE_UP_DTYPE)
            u["var"] = np.arange(c.var_count)
            d = self.rng.random(c.var_count) < 0.5
            u["is_def"] = d
            u["is_true"] = d & ((self.rng.random(c.var_count) < c.agree) == self.sigma)
            loads = np.zeros(WORDS, LOAD_DTYPE)
            loads["word"] = np.arange(WORDS)
            submit(t, np.zeros(WORDS, np.uint64), base_ups=u, loads=loads)
            t.step_wait()
            best, ms = t.autotune(np.full(WORDS, ONES))
        finally:
            t.close()
        self.srv.set_tuning(best)
        self.log(f"tuned on {n} clauses: {best} ({ms:.3f} ms per sweep)")
        return best

    def _update_base(self, u: np.ndarray) -> None:
        self.srv.update_base(u)
        if self.keep_history:
            self.base_history.append(u.copy())

    def initial_bases(self) -> None:
        """Every solver's base: each variable assigned with the solver's
        probability, planted polarity (uniform without the planted
        assignment); published, then checked by a first empty-patch request."""
        c, rng = self.cfg, self.rng
        for s in range(c.solvers):
            vs = np.flatnonzero(rng.random(c.var_count) < self.p_assigned[s]).astype(np.uint32)
            u = np.zeros(vs.size, BASE_UP_DTYPE)
            u["var"] = vs
            u["base"] = s
            u["is_def"] = 1
            agree = c.agree if c.planted else 0.5
            u["is_true"] = (rng.random(vs.size) < agree) == self.sigma[vs]
            self._update_base(u)
        rid = self.srv.submit(0, 64, np.zeros(0, PATCH_DTYPE))
        while rid not in self.srv.poll(16, 1_000_000):
            pass
        self.srv.release(rid)

    def schedule(self) -> list[Request]:
        c, rng = self.cfg, self.rng
        n = rng.poisson(c.rate * c.duration)
        t = np.sort(rng.random(n) * c.duration)
        return [Request(i, float(t[i]), int(rng.integers(0, c.solvers)),
                        int(rng.integers(c.set_min, c.set_max + 1)),
                        int(rng.integers(0, c.pool))) for i in range(n)]

    def _clauses(self, n: int):
        c = self.cfg
        lo = self.cl_next
        hi = min(lo + n, c.clause_pool)
        self.cl_next = hi if hi < c.clause_pool else 0
        a, b = self.cl_off[lo], self.cl_off[hi]
        return self.cl_lits[a:b], self.cl_lens[lo:hi], lo, hi

    def _drift(self, s: int) -> np.ndarray:
        c, rng = self.cfg, self.rng
        u = np.zeros(c.drift, BASE_UP_DTYPE)
        u["var"] = rng.integers(0, c.var_count, size=c.drift)
        u["base"] = s
        d = rng.random(c.drift) < self.p_assigned[s]
        u["is_def"] = d
        agree = c.agree if c.planted else 0.5
        u["is_true"] = d & ((rng.random(c.drift) < agree) == self.sigma[u["var"]])
        return u

    def preload(self, n: int, chunk: int = 1_000_000) -> int:
        """Adds n clauses of the workload's clause stream before the run (in
        chunks, each appended by the step of one one-assignment request);
        returns the clauses in the DB."""
        left = n
        while left > 0:
            k = min(left, chunk)
            added = 0
            while added < k:
                lits, lens, lo, hi = self._clauses(k - added)
                first = self.srv.add_clauses(lits, lens)
                if first < 0:  # the budget is full
                    return self.srv.stats()["clauses"]
                if self.keep_history:
                    self.clause_ids.append((first, lo, hi))
                added += hi - lo
            left -= k
            rid = self.srv.submit(0, 1, np.zeros(0, PATCH_DTYPE))
            if rid is None:
                raise RuntimeError("the request ring is full")
            done = False
            while not done:
                for i in self.srv.poll(16, 1_000_000):
                    self.srv.result(i, records=False)
                    self.srv.release(i)
                    done |= i == rid
        return self.srv.stats()["clauses"]

    def request_patches(self, r: Request) -> np.ndarray:
        nw = (r.k + 63) // 64
        return self.pool[r.payload][:self.pool_off[r.payload][nw]]

    def run(self, keep=None) -> ServeResult:
        """keep(request) -> bool selects requests whose records are kept."""
        c, srv = self.cfg, self.srv
        self.initial_bases()
        reqs = self.schedule()
        self.log(f"{len(reqs)} requests over {c.duration:.0f} s")
        by_id: dict[int, Request] = {}
        t_sub: dict[int, float] = {}
        lat, arrive, lags = [], [], []
        kept = {}
        st0 = srv.stats()
        ai = 0
        max_queue = 0
        last_cl = 0.0
        last_log = 0.0
        backlog: list[Request] = []
        st = srv.stats()
        samples = [(0.0, st["steps"], st["device_ms"], st["clauses"])]
        start = time.perf_counter()

        def collect(ids):
            for rid in ids:
                r = by_id.pop(rid)
                want = keep is not None and keep(r)
                meta, recs = srv.result(rid, records=want)
                srv.release(rid)
                lag = (t_sub.pop(rid) - r.t_arrive) * 1e3
                lags.append(lag)
                lat.append(lag + meta["latency_ms"])
                arrive.append(r.t_arrive)
                if want:
                    kept[rid] = (r, meta, recs)

        while ai < len(reqs) or by_id or backlog:
            now = time.perf_counter() - start
            while ai < len(reqs) and reqs[ai].t_arrive <= now:
                r = reqs[ai]
                ai += 1
                self._update_base(self._drift(r.solver))
                if now <= c.duration:
                    n_cl = int(c.clause_rate * (now - last_cl))
                    last_cl = now
                    if n_cl:
                        lits, lens, lo, hi = self._clauses(n_cl)
                        first = srv.add_clauses(lits, lens)
                        if first >= 0 and self.keep_history:
                            self.clause_ids.append((first, lo, hi))
                backlog.append(r)
            while backlog:
                r = backlog[0]
                t = time.perf_counter() - start
                rid = srv.submit(r.solver, r.k, self.request_patches(r))
                if rid is None:
                    break  # ring full: retry after completions
                backlog.pop(0)
                by_id[rid] = r
                t_sub[rid] = t
            max_queue = max(max_queue, len(by_id) + len(backlog))
            nxt = reqs[ai].t_arrive - (time.perf_counter() - start) if ai < len(reqs) else 0.002
            collect(srv.poll(1024, int(max(0.0, min(nxt, 0.002)) * 1e6)))
            if now - last_log > 5.0:
                last_log = now
                st = srv.stats()
                samples.append((now, st["steps"], st["device_ms"], st["clauses"]))
                dms = 0.0
                if len(samples) > 1 and samples[-1][1] > samples[-2][1]:
                    dms = (samples[-1][2] - samples[-2][2]) / (samples[-1][1] - samples[-2][1])
                self.log(f"t={now:.1f}s steps {st['steps']} served {len(lat)} queue "
                         f"{len(by_id) + len(backlog)} DB {st['clauses']} clauses, "
                         f"{dms:.3f} device ms per step")
        elapsed = time.perf_counter() - start
        st = srv.stats()
        return ServeResult(len(lat), st["steps"] - st0["steps"], elapsed,
                           st["checks"] - st0["checks"], np.array(lat), st["clauses"],
                           st["literals"], st["records"] - st0["records"], max_queue,
                           st["device_ms"] - st0["device_ms"], np.array(arrive), kept,
                           samples + [(elapsed, st["steps"], st["device_ms"], st["clauses"])],
                           st["host_grouped"] - st0["host_grouped"],
                           st["graph_captures"] - st0["graph_captures"],
                           {"submit_ms_mean": (st["submit_ms_total"] - st0["submit_ms_total"])
                            / max(1, st["steps"] - st0["steps"]),
                            "submit_ms_max": st["submit_ms_max"],
                            "wait_ms_max": st["wait_ms_max"]}, np.array(lags),
                           srv.step_stats())

    def close(self):
        self.srv.close()
