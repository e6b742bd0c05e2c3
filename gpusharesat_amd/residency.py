"""Python handle on the clause residency scorer (include/gss_residency.h) and
the residency simulation used by the benchmark and the tests.

The simulation follows the learned-residency workload: clauses arrive every
round into a GPU budget; each round the 1024 assignments of 16 solvers (one
64-slot word each) are checked against the whole DB; when the DB exceeds its
budget the policy removes about half of the clauses of length >= 3.  The
figure of merit is hit pairs (clause, assignment) per GB of clause literals
checked.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib
from .engine import Engine, VAR_UPDATE_DTYPE

FEATURES = 12
HIDDEN = 64
PARAMS = HIDDEN * FEATURES + HIDDEN + HIDDEN * HIDDEN + HIDDEN + HIDDEN + 1
POLICY_ACTIVITY = 0
POLICY_LEARNED = 1


def _ptr(a: np.ndarray):
    return a.ctypes.data if a is not None and a.size else None


class Residency:
    def __init__(self, engine: Engine, max_tags: int, decay: float = 0.99999, seed: int = 1):
        self.engine = engine
        self.max_tags = max_tags
        self._h = ctypes.c_void_p()
        _lib.call("gss_res_create", engine.handle, max_tags, decay, seed, ctypes.byref(self._h))
        engine.add_dependent(self)  # the scorer goes before its engine

    def close(self) -> None:
        if self._h:
            _lib.call("gss_res_destroy", self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, lits, lens, tags, round_: int) -> None:
        lits = np.ascontiguousarray(lits, dtype=np.uint32)
        lens = np.ascontiguousarray(lens, dtype=np.int32)
        tags = np.ascontiguousarray(tags, dtype=np.uint32)
        _lib.call("gss_res_add", self._h, lens.size, _ptr(lits), _ptr(lens), _ptr(tags), round_)

    def account(self, round_: int) -> int:
        n = ctypes.c_int64()
        _lib.call("gss_res_account", self._h, round_, ctypes.byref(n))
        return n.value

    def snapshot(self, round_: int, tags) -> None:
        tags = np.ascontiguousarray(tags, dtype=np.uint32)
        _lib.call("gss_res_snapshot", self._h, round_, tags.size, _ptr(tags))

    def train(self, round_: int, horizon: int = 5, epochs: int = 2, batch: int = 2048,
              lr: float = 0.02) -> tuple[float, int]:
        loss, n = ctypes.c_float(), ctypes.c_int64()
        _lib.call("gss_res_train", self._h, round_, horizon, epochs, batch, lr,
                  ctypes.byref(loss), ctypes.byref(n))
        return loss.value, n.value

    def reduce(self, policy: int, round_: int) -> tuple[np.ndarray, int]:
        ev = np.zeros(self.max_tags, dtype=np.uint8)
        kept = ctypes.c_int64()
        _lib.call("gss_res_reduce", self._h, policy, round_, _ptr(ev), ctypes.byref(kept))
        return ev.astype(bool), kept.value

    def features(self, round_: int, tags) -> np.ndarray:
        tags = np.ascontiguousarray(tags, dtype=np.uint32)
        out = np.empty((tags.size, FEATURES), dtype=np.float32)
        _lib.call("gss_res_features", self._h, round_, tags.size, _ptr(tags), _ptr(out))
        return out

    def activity(self, tags) -> np.ndarray:
        tags = np.ascontiguousarray(tags, dtype=np.uint32)
        out = np.empty(tags.size, dtype=np.float32)
        _lib.call("gss_res_activity", self._h, tags.size, _ptr(tags), _ptr(out))
        return out

    @property
    def params(self) -> np.ndarray:
        out = np.empty(PARAMS, dtype=np.float32)
        _lib.call("gss_res_get_params", self._h, _ptr(out))
        return out

    @params.setter
    def params(self, p) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != PARAMS:
            raise ValueError(f"expected {PARAMS} parameters")
        _lib.call("gss_res_set_params", self._h, _ptr(p))

    def grad(self, x, y, wt=None) -> tuple[float, np.ndarray]:
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        wt = None if wt is None else np.ascontiguousarray(wt, dtype=np.float32)
        g = np.empty(PARAMS, dtype=np.float32)
        loss = ctypes.c_float()
        _lib.call("gss_res_grad", self._h, y.size, _ptr(x), _ptr(y), _ptr(wt),
                  ctypes.byref(loss), _ptr(g))
        return loss.value, g

    def score(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32)
        out = np.empty(x.shape[0], dtype=np.float32)
        _lib.call("gss_res_score", self._h, x.shape[0], _ptr(x), _ptr(out))
        return out


# ---- the drifting hot-set workload ---------------------------------------------

@dataclass
class SimConfig:
    var_count: int = 2_000_000
    solvers: int = 16                # one 64-slot word each: 1024 assignments
    hot_vars: int = 5000             # per solver
    redraw_every: int = 50           # rounds
    arrivals: int = 60_000_000       # clauses over the whole run
    rounds: int = 300
    budget: int = 20_000_000         # clauses on the GPU
    hot_fraction: float = 0.5        # arriving clauses built on a current hot set
    sigma_false: float = 0.9         # hot-clause literal false under the solver's polarity
    p_assigned: float = 0.9          # hot variable assigned in the round's base
    agree: float = 0.98              # assigned hot variable takes the solver's polarity
    slot_flips: int = 8              # hot variables made undefined per slot
    seed: int = 11


def _positions(off: np.ndarray, idx: np.ndarray, lens: np.ndarray) -> np.ndarray:
    """Literal positions of clauses idx (in order) in a CSR layout."""
    ln = lens[idx].astype(np.int64)
    starts = np.repeat(off[idx], ln)
    first = np.zeros(idx.size, dtype=np.int64)
    np.cumsum(ln[:-1], out=first[1:])
    return starts + np.arange(int(ln.sum())) - np.repeat(first, ln)


class DriftingHotSets:
    """Clause arrivals and per-round assignments with per-solver hot sets.

    Solver s has a hot set of `hot_vars` variables with a polarity sigma_s,
    redrawn every `redraw_every` rounds.  Each round it assigns a base over its
    hot set (assigned w.p. p_assigned, sigma_s's value w.p. agree) and its 64
    slots are that base with `slot_flips` hot variables left undefined; cold
    variables are undefined.  A fraction `hot_fraction` of arriving clauses is
    built on a random solver's current hot set with each literal false under
    sigma_s w.p. sigma_false; the others are uniform over all variables.
    Lengths are 2 + Geometric(0.1) capped at 200, variables distinct.
    """

    def __init__(self, cfg: SimConfig):
        self.cfg = cfg
        self.rng = np.random.default_rng(cfg.seed)
        self.words = cfg.solvers
        self.hot = np.zeros((cfg.solvers, cfg.hot_vars), dtype=np.int64)
        self.sigma = np.zeros((cfg.solvers, cfg.hot_vars), dtype=bool)
        self.per_round = cfg.arrivals // cfg.rounds
        self._prev_hot = None

    def _redraw(self):
        c = self.cfg
        self._prev_hot = self.hot.copy()
        for s in range(c.solvers):
            self.hot[s] = self.rng.choice(c.var_count, size=c.hot_vars, replace=False)
        self.sigma = self.rng.random(self.hot.shape) < 0.5

    def clauses(self, n: int) -> tuple[np.ndarray, np.ndarray]:
        from . import workloads
        c, rng = self.cfg, self.rng
        lens = workloads.clause_lengths(rng, n, 0.1, 200)
        hot = rng.random(n) < c.hot_fraction
        lens[hot] = np.minimum(lens[hot], c.hot_vars)
        lits = np.empty(int(lens.sum()), dtype=np.uint32)
        off = np.zeros(n + 1, dtype=np.int64)
        np.cumsum(lens, out=off[1:])
        idx_h = np.flatnonzero(hot)
        idx_c = np.flatnonzero(~hot)
        if idx_c.size:
            v = workloads.distinct_vars(rng, lens[idx_c], c.var_count)
            sign = rng.integers(0, 2, size=v.size, dtype=np.uint32)
            lits[_positions(off, idx_c, lens)] = (v.astype(np.uint32) << np.uint32(1)) | sign
        if idx_h.size:
            k = workloads.distinct_vars(rng, lens[idx_h], c.hot_vars)  # index into a hot set
            owner = np.repeat(rng.integers(0, c.solvers, size=idx_h.size), lens[idx_h])
            v = self.hot[owner, k]
            sig = self.sigma[owner, k]
            # literal false under sigma: positive literal of a false variable
            # (sign 0) or negative literal of a true one (sign 1)
            false_lit = rng.random(v.size) < c.sigma_false
            sign = np.where(false_lit, sig, ~sig).astype(np.uint32)
            lits[_positions(off, idx_h, lens)] = (v.astype(np.uint32) << np.uint32(1)) | sign
        return lits, lens

    def updates(self, rnd: int) -> np.ndarray:
        """Table updates of round rnd (one per (var, word) touched)."""
        c, rng = self.cfg, self.rng
        parts = []
        if rnd % c.redraw_every == 0:
            self._redraw()
            if rnd > 0:
                for s in range(c.solvers):  # the old hot set becomes undefined
                    u = np.zeros(c.hot_vars, VAR_UPDATE_DTYPE)
                    u["var"] = self._prev_hot[s]
                    u["word"] = s
                    parts.append(u)
        ones = np.uint64(0xFFFFFFFFFFFFFFFF)
        for s in range(c.solvers):
            assigned = rng.random(c.hot_vars) < c.p_assigned
            val = np.where(rng.random(c.hot_vars) < c.agree, self.sigma[s], ~self.sigma[s])
            d = np.where(assigned, ones, np.uint64(0))
            # slot_flips undefined hot variables per slot
            fv = rng.integers(0, c.hot_vars, size=(64, c.slot_flips))
            clear = np.zeros(c.hot_vars, dtype=np.uint64)
            np.bitwise_or.at(clear, fv.ravel(),
                             np.repeat(np.uint64(1) << np.arange(64, dtype=np.uint64),
                                       c.slot_flips))
            d &= ~clear
            t = np.where(val, d, np.uint64(0))
            u = np.zeros(c.hot_vars, VAR_UPDATE_DTYPE)
            u["var"] = self.hot[s]
            u["word"] = s
            u["is_def"] = d
            u["is_true"] = t
            parts.append(u)
        ups = np.concatenate(parts)
        # one update per (var, word): a var redrawn into the same solver keeps the new one
        key = ups["var"].astype(np.int64) * 64 + ups["word"]
        _, last = np.unique(key[::-1], return_index=True)
        return ups[np.sort(ups.size - 1 - last)]


@dataclass
class SimResult:
    policy: int
    hit_pairs: int
    bytes_checked: int
    rounds: int
    reductions: int
    db_end: int
    train_loss: float

    @property
    def hits_per_gb(self) -> float:
        return self.hit_pairs / (self.bytes_checked / 1e9)


def simulate(cfg: SimConfig, policy: int, device: int = 0, train: bool = True,
             snapshot: int = 8192, horizon: int = 5, log=None, after_reduce=None) -> SimResult:
    """Runs the residency workload under one policy (the learned one trains
    online whatever the policy, so both runs do the same work).
    after_reduce(engine, round, evicted_tags), if given, runs after every
    reduction (the tests check the kept DB there)."""
    e = Engine(device)
    try:
        return _simulate_on(e, cfg, policy, train, snapshot, horizon, log, after_reduce)
    finally:
        e.close()  # closes the scorer first, also when the run raised


def _simulate_on(e: Engine, cfg: SimConfig, policy: int, train: bool, snapshot: int,
                 horizon: int, log, after_reduce) -> SimResult:
    gen = DriftingHotSets(cfg)
    e.enable_tags()
    e.reset_assignments(cfg.var_count, gen.words)
    max_tags = cfg.budget + cfg.arrivals // cfg.rounds + 1024
    res = Residency(e, max_tags, seed=cfg.seed)
    free = np.arange(max_tags, dtype=np.uint32)[::-1].copy()  # stack of unused tags
    top = max_tags
    alive = np.zeros(max_tags, dtype=bool)
    valid = np.full(gen.words, np.uint64(0xFFFFFFFFFFFFFFFF))
    rng = np.random.default_rng(cfg.seed + 1)
    pairs = checked = reductions = 0
    loss = 0.0
    for rnd in range(cfg.rounds):
        lits, lens = gen.clauses(gen.per_round)
        tags = free[top - lens.size:top][::-1].copy()
        top -= lens.size
        res.add(lits, lens, tags, rnd)
        alive[tags] = True
        e.apply(gen.updates(rnd))
        checked += 4 * e.totals()[1]
        e.sweep(valid)
        pairs += res.account(rnd)
        live = np.flatnonzero(alive)
        if train and live.size:
            res.snapshot(rnd, rng.choice(live, size=min(snapshot, live.size), replace=False))
            l, n = res.train(rnd, horizon=horizon)
            if n:
                loss = l
        if e.totals()[0] > cfg.budget:
            ev, kept = res.reduce(policy, rnd)
            gone = np.flatnonzero(ev).astype(np.uint32)
            alive[gone] = False
            free[top:top + gone.size] = gone
            top += gone.size
            reductions += 1
            if after_reduce is not None:
                after_reduce(e, rnd, gone)
        if log and rnd % 25 == 0:
            log(f"policy {policy} round {rnd}: DB {e.totals()[0]}, hit pairs {pairs}, "
                f"loss {loss:.4f}")
    out = SimResult(policy, pairs, checked, cfg.rounds, reductions, e.totals()[0], loss)
    res.close()
    return out
