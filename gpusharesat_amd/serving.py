"""Serving loop over the batched step: requests (sets of assignments from a
solver) arrive continuously and are admitted into the next step as table
words free up, clause inserts are admitted under an HBM budget, two steps
are in flight, and every request's records reach the host as soon as its
step is done.

A request of k assignments from solver s takes ceil(k / 64) consecutive
table words.  The step loads those words from solver s's base (device-side
packed planes, gss_bases_reset / gss_word_load), applies the base drift since
s's previous request first (gss_base_update), and then the request's own
per-slot differences as masked patches (gss_var_patch).  So the host sends
only differences, never whole assignments, and no request depends on which
words it lands in: its records are the same alone or in any mix.

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
from collections import deque
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
                ("records", ctypes.c_int32), ("reserved", ctypes.c_int32)]


def _p(a):
    return a.ctypes.data if a is not None and a.size else None


RECORDS_FULL = 0
RECORDS_IDS = 1


def submit(e: Engine, valid, lits=None, lens=None, patches=None, base_ups=None, loads=None,
           ups=None, last_pos=None, records: int = RECORDS_FULL) -> None:
    """gss_step_submit with numpy arrays (kept alive by the caller until the
    call returns; the engine copies them into its pinned staging)."""
    d = _StepDesc()
    d.records = records
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
    pool: int = 48                  # distinct request payloads per size
    clause_pool: int = 2_000_000
    max_inflight: int = 2           # steps on the device at once (1: each step starts once the previous has left it)
    seed: int = 5


@dataclass
class Request:
    idx: int
    t_arrive: float
    solver: int
    k: int
    payload: int
    words: tuple = ()
    t_done: float = 0.0
    records: int = 0


@dataclass
class ServeResult:
    requests: int
    steps: int
    elapsed: float
    checks: int
    latency_ms: np.ndarray
    clauses_end: int
    literals_end: int
    records: int
    max_queue: int
    step_ms: list = field(default_factory=list)
    arrive_s: np.ndarray = None  # arrival time of the request of each latency

    @property
    def checks_per_s(self) -> float:
        return self.checks / self.elapsed


class Server:
    """The serving loop (host side) over one engine."""

    def __init__(self, cfg: ServeConfig, device: int = 0, log=None):
        from . import workloads
        self.cfg = cfg
        self.log = log or (lambda m: None)
        rng = np.random.default_rng(cfg.seed)
        self.rng = rng
        self.e = Engine(device)
        self.e.reset_assignments(cfg.var_count, WORDS)
        bases_reset(self.e, cfg.solvers)
        self.p_assigned = rng.uniform(cfg.p_min, cfg.p_max, size=cfg.solvers)
        self.sigma = rng.random(cfg.var_count) < 0.5  # the planted assignment
        # payload pools: per solver-independent sets of 1024 slots; a request
        # of k assignments takes the patches of its first ceil(k/64) words and
        # the slots below k
        # kept as uint64 [n, 4] rows sorted by word (first column = var |
        # word << 32), with the first row of every word, so a request takes a
        # prefix and shifts its words with one add
        self.pool, self.pool_off = [], []
        for _ in range(cfg.pool):
            pt = slot_patches(rng, cfg.var_count, 1024, cfg.flips, 0.5, self.sigma, cfg.agree)
            pt = pt[np.argsort(pt["word"], kind="stable")]
            self.pool.append(pt.view(np.uint64).reshape(-1, 4).copy())
            self.pool_off.append(np.searchsorted(pt["word"], np.arange(17)))
        self.log(f"payload pool: {sum(p.shape[0] for p in self.pool)} patches")
        lens = workloads.clause_lengths(rng, cfg.clause_pool, 0.1, 200)
        self.cl_lens = lens
        self.cl_off = np.zeros(lens.size + 1, np.int64)
        np.cumsum(lens, out=self.cl_off[1:])
        v = workloads.distinct_vars(rng, lens, cfg.var_count)
        st = rng.random(v.size) < cfg.lit_true
        # literal 2v + 1 is the negation: true under sigma iff sigma(v) is false
        self.cl_lits = ((v << 1) | (st != self.sigma[v])).astype(np.uint32)
        self.cl_next = 0

    def initial_bases(self) -> None:
        """Every solver's base: each variable assigned with the solver's
        probability, polarity Bernoulli(0.5), sent in chunks of steps."""
        c, rng = self.cfg, self.rng
        for s in range(c.solvers):
            vs = np.flatnonzero(rng.random(c.var_count) < self.p_assigned[s]).astype(np.uint32)
            u = np.zeros(vs.size, BASE_UP_DTYPE)
            u["var"] = vs
            u["base"] = s
            u["is_def"] = 1
            u["is_true"] = (rng.random(vs.size) < c.agree) == self.sigma[vs]
            submit(self.e, np.zeros(WORDS, np.uint64), base_ups=u)
            self.e.step_wait()

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
        return self.cl_lits[a:b], self.cl_lens[lo:hi]

    def _drift(self, s: int) -> np.ndarray:
        c, rng = self.cfg, self.rng
        u = np.zeros(c.drift, BASE_UP_DTYPE)
        u["var"] = rng.integers(0, c.var_count, size=c.drift)  # repeats removed in build()
        u["base"] = s
        d = rng.random(c.drift) < self.p_assigned[s]
        u["is_def"] = d
        u["is_true"] = d & ((rng.random(c.drift) < c.agree) == self.sigma[u["var"]])
        return u

    def build(self, reqs: list[Request], n_clauses: int):
        """Step inputs for admitted requests (words assigned in order)."""
        valid = np.zeros(WORDS, np.uint64)
        loads, patches, bups = [], [], []
        w0 = 0
        for r in reqs:
            nw = (r.k + 63) // 64
            r.words = (w0, w0 + nw)
            bups.append(self._drift(r.solver))
            ld = np.zeros(nw, LOAD_DTYPE)
            ld["word"] = np.arange(w0, w0 + nw)
            ld["base"] = r.solver
            loads.append(ld)
            pool, off = self.pool[r.payload], self.pool_off[r.payload]
            last = r.k - 64 * (nw - 1)  # slots used in the request's last word
            full_end = off[nw - 1] if last < 64 else off[nw]
            patches.append((pool[:full_end], w0))
            if last < 64:  # the last word keeps only the request's slots
                t = pool[off[nw - 1]:off[nw]].copy()
                t[:, 1] &= np.uint64((1 << last) - 1)
                patches.append((t[t[:, 1] != 0], w0))
            valid[w0:w0 + nw - 1] = ONES
            valid[w0 + nw - 1] = ONES if last == 64 else np.uint64((1 << last) - 1)
            w0 += nw
        lits, lens = self._clauses(n_clauses) if n_clauses else (None, None)
        cat = lambda xs, dt: np.concatenate(xs) if xs else np.zeros(0, dt)
        bu = cat(bups, BASE_UP_DTYPE)
        if bu.size:  # one update per (base, var) in a step: the last one wins
            key = (bu["base"].astype(np.int64) << 32) | bu["var"]
            _, last = np.unique(key[::-1], return_index=True)
            bu = bu[bu.size - 1 - last]
        # one pass: copy every request's rows into place and shift its words
        out = np.empty((sum(a.shape[0] for a, _ in patches), 4), np.uint64)
        at = 0
        for a, w0 in patches:
            np.add(a, np.array([w0 << 32, 0, 0, 0], np.uint64), out=out[at:at + a.shape[0]])
            at += a.shape[0]
        pt = out.view(PATCH_DTYPE).ravel()
        return dict(valid=valid, lits=lits, lens=lens, patches=pt, base_ups=bu,
                    loads=cat(loads, LOAD_DTYPE))

    def tune(self, clauses: int = 2_000_000) -> dict:
        """Chooses the sweep tiles on a scratch engine holding `clauses` pooled
        clauses and a 64-word table of the solvers' bases, and applies them
        to the serving engine (whose DB starts empty)."""
        c = self.cfg
        t = Engine(self.e.device)
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
        self.e.apply_tuning(best)
        self.log(f"tuned on {n} clauses: {best} ({ms:.3f} ms per sweep)")
        return best

    def run(self) -> ServeResult:
        c, e = self.cfg, self.e
        self.initial_bases()
        reqs = self.schedule()
        self.log(f"{len(reqs)} requests over {c.duration:.0f} s")
        pending: deque[int] = deque()
        inflight: deque = deque()
        lat = []
        arrive = []
        ai = 0
        steps = 0
        checks = 0
        records = 0
        max_queue = 0
        step_ms = []
        lits_db = e.totals()[1]
        start = time.perf_counter()
        last_cl = 0.0
        last_log = 0.0

        def arrivals(now):
            nonlocal ai, max_queue
            while ai < len(reqs) and reqs[ai].t_arrive <= now:
                pending.append(ai)
                ai += 1
            max_queue = max(max_queue, len(pending))

        def admit(now):
            """Submits one step with the pending requests whose words fit."""
            nonlocal last_cl, lits_db, checks, steps, last_log
            batch, used = [], 0
            while pending:
                r = reqs[pending[0]]
                nw = (r.k + 63) // 64
                if used + nw > WORDS:
                    break
                batch.append(r)
                used += nw
                pending.popleft()
            n_cl = 0
            if now <= c.duration:
                n_cl = int(c.clause_rate * (now - last_cl))
                if 4 * (lits_db + 11 * n_cl) > c.hbm_budget:
                    n_cl = 0
                last_cl = now
            st = self.build(batch, n_cl)
            submit(e, **st)
            if st["lens"] is not None:
                lits_db += int(st["lits"].size)
            checks += lits_db * sum(r.k for r in batch)
            inflight.append(([r.idx for r in batch], now))
            steps += 1
            if now - last_log > 5.0:
                last_log = now
                self.log(f"t={now:.1f}s steps {steps} served {len(lat)} queue {len(pending)} "
                         f"DB {e.totals()[0]} clauses")

        def finish():
            """Waits for the oldest step; while its records travel to the host
            the next step is admitted, then the records are read."""
            nonlocal records
            ids, t_sub = inflight.popleft()
            _, ms = e.step_wait()
            now = time.perf_counter() - start
            arrivals(now)
            if pending:
                admit(now)
            h = e.step_hits_view()
            t = time.perf_counter() - start
            records += h.size
            step_ms.append(ms)
            for i in ids:
                r = reqs[i]
                r.t_done = t
                lat.append(t - r.t_arrive)
                arrive.append(r.t_arrive)

        while True:
            now = time.perf_counter() - start
            arrivals(now)
            if ai >= len(reqs) and not pending and not inflight:
                break
            if len(inflight) >= c.max_inflight or (inflight and not pending):
                finish()
                continue
            if not pending:
                nxt = reqs[ai].t_arrive if ai < len(reqs) else now
                time.sleep(max(0.0, min(nxt - now, 0.0005)))
                continue
            admit(now)
        elapsed = time.perf_counter() - start
        cl, li = e.totals()
        return ServeResult(len(lat), steps, elapsed, checks, np.array(lat) * 1e3, cl, li,
                           records, max_queue, step_ms, np.array(arrive))

    def close(self):
        self.e.close()
