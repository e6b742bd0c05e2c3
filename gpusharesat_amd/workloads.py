"""Seeded synthetic workloads of the clause-vs-assignment check.

Two instance shapes are used throughout the tests and the benchmark:

  plain-64:   10k vars, 100k clauses, length 2 + Geometric(0.15) capped at 100,
              64 assignments each setting every var with probability 0.6,
              polarity Bernoulli(0.5); 1% of clauses planted falsified and 1%
              planted unit under a random assignment.
  sweep-1024: 1M vars, 4M clauses, length 2 + Geometric(0.1) capped at 200,
              1024 assignments, each var assigned with probability 0.4,
              polarity Bernoulli(0.5); 0.1% planted hits.

Geometric(p) counts failures before the first success (support 0, 1, ...), so
the shortest clause has 2 literals.  Clause variables are distinct within a
clause and uniform; signs are uniform.  Literals are packed 2*var + sign
(sign 1 = negative).  Assignments are var-major uint64 planes is_def/is_true
[vars, words]; bit b of word w is assignment 64*w + b.  The generator is
numpy's PCG64 seeded with the given seed.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass
class Instance:
    var_count: int
    lits: np.ndarray          # uint32, concatenated
    lens: np.ndarray          # int32 per clause
    is_def: np.ndarray        # uint64 [vars, words]
    is_true: np.ndarray       # uint64 [vars, words]
    assignments: int
    planted: dict = field(default_factory=dict)

    @property
    def offsets(self) -> np.ndarray:
        off = np.zeros(self.lens.size + 1, dtype=np.int64)
        np.cumsum(self.lens, out=off[1:])
        return off

    @property
    def words(self) -> int:
        return self.is_def.shape[1]

    @property
    def valid(self) -> np.ndarray:
        v = np.zeros(self.words, dtype=np.uint64)
        full, rem = divmod(self.assignments, 64)
        v[:full] = np.uint64(0xFFFFFFFFFFFFFFFF)
        if rem:
            v[full] = np.uint64((1 << rem) - 1)
        return v

    @property
    def used_vars(self) -> int:
        return int(np.unique(self.lits >> np.uint32(1)).size)


def clause_lengths(rng: np.random.Generator, n: int, p: float, cap: int) -> np.ndarray:
    lens = 2 + (rng.geometric(p, size=n) - 1)
    return np.minimum(lens, cap).astype(np.int32)


def distinct_vars(rng: np.random.Generator, lens: np.ndarray, var_count: int) -> np.ndarray:
    """Uniform variables, distinct within each clause: clauses of one length
    form a matrix whose rows are redrawn until no row repeats a variable."""
    if lens.size and int(lens.max()) > var_count:
        raise ValueError("clause longer than the variable count")
    total = int(lens.sum(dtype=np.int64))
    off = np.zeros(lens.size + 1, dtype=np.int64)
    np.cumsum(lens, out=off[1:])
    out = np.empty(total, dtype=np.int64)
    order = np.argsort(lens, kind="stable")
    sorted_lens = lens[order]
    bounds = np.flatnonzero(np.diff(sorted_lens)) + 1
    for grp in np.split(order, bounds):
        if grp.size == 0:
            continue
        L = int(lens[grp[0]])
        idx = off[grp][:, None] + np.arange(L)[None, :]
        if L * L > 2 * var_count:
            # redraws would rarely succeed: take the first L of a random
            # permutation of the variables, per row
            rows = max(1, (1 << 24) // var_count)
            m = np.empty((grp.size, L), dtype=np.int64)
            for r0 in range(0, grp.size, rows):
                keys = rng.random((min(rows, grp.size - r0), var_count))
                m[r0:r0 + rows] = np.argpartition(keys, L - 1, axis=1)[:, :L]
            out[idx] = m
            continue
        m = rng.integers(0, var_count, size=(grp.size, L), dtype=np.int64)
        bad = np.arange(grp.size)
        for _ in range(10_000):
            s = np.sort(m[bad], axis=1)
            dup = (s[:, 1:] == s[:, :-1]).any(axis=1) if L > 1 else np.zeros(bad.size, bool)
            bad = bad[dup]
            if bad.size == 0:
                break
            m[bad] = rng.integers(0, var_count, size=(bad.size, L), dtype=np.int64)
        else:
            raise RuntimeError("could not make clause variables distinct")
        out[idx] = m
    return out


def bernoulli_words(rng: np.random.Generator, var_count: int, assignments: int,
                    p: float, block: int = 1 << 15) -> np.ndarray:
    """uint64 [vars, ceil(A/64)] with each of the first A bits set with prob p."""
    words = (assignments + 63) // 64
    out = np.zeros((var_count, words), dtype=np.uint64)
    thr = np.uint16(min(65535, round(p * 65536)))
    for v0 in range(0, var_count, block):
        nv = min(block, var_count - v0)
        raw = rng.bit_generator.random_raw(nv * words * 16).astype(np.uint64, copy=False)
        bits = raw.view(np.uint16).reshape(nv, words * 64) < thr
        bits[:, assignments:] = False
        packed = np.packbits(bits, axis=1, bitorder="little")
        out[v0:v0 + nv] = packed.view(np.uint64).reshape(nv, words)
    return out


def random_assignments(rng: np.random.Generator, var_count: int, assignments: int,
                       p_assigned: float) -> tuple[np.ndarray, np.ndarray]:
    is_def = bernoulli_words(rng, var_count, assignments, p_assigned)
    is_true = bernoulli_words(rng, var_count, assignments, 0.5) & is_def
    return is_def, is_true


def _bit(plane: np.ndarray, var: np.ndarray, a: np.ndarray) -> np.ndarray:
    w = (a // 64).astype(np.int64)
    b = (a % 64).astype(np.uint64)
    return ((plane[var, w] >> b) & np.uint64(1)).astype(bool)


def plant(rng: np.random.Generator, inst: Instance, frac_false: float, frac_unit: float) -> None:
    """Rewrites a random set of clauses so that each is falsified (or unit)
    under one randomly chosen assignment."""
    n = inst.lens.size
    nf, nu = int(round(n * frac_false)), int(round(n * frac_unit))
    chosen = rng.choice(n, size=nf + nu, replace=False)
    off = inst.offsets
    V = inst.var_count
    kinds = np.array([0] * nf + [1] * nu)
    assig = rng.integers(0, inst.assignments, size=chosen.size)
    for c, kind, a in zip(chosen, kinds, assig):
        L = int(inst.lens[c])
        want_def = np.ones(L, dtype=bool)
        if kind == 1:
            want_def[rng.integers(0, L)] = False
        vs = np.empty(L, dtype=np.int64)
        used = set()
        for i in range(L):
            while True:
                v = int(rng.integers(0, V))
                if v in used:
                    continue
                if _bit(inst.is_def, np.array([v]), np.array([a]))[0] == want_def[i]:
                    break
            used.add(v)
            vs[i] = v
        tru = _bit(inst.is_true, vs, np.full(L, a)) & want_def
        # literal false <=> (positive and var false) or (negative and var true)
        sign = np.where(want_def, tru, rng.integers(0, 2, size=L).astype(bool))
        inst.lits[off[c]:off[c] + L] = (vs.astype(np.uint32) << np.uint32(1)) | sign.astype(np.uint32)
    inst.planted = {"clauses": chosen, "kinds": kinds, "assignments": assig}


def make_instance(seed: int, var_count: int, clause_count: int, p_len: float, cap: int,
                  assignments: int, p_assigned: float, frac_false: float = 0.0,
                  frac_unit: float = 0.0) -> Instance:
    rng = np.random.default_rng(seed)
    lens = clause_lengths(rng, clause_count, p_len, cap)
    vars_ = distinct_vars(rng, lens, var_count)
    signs = rng.integers(0, 2, size=vars_.size, dtype=np.uint32)
    lits = (vars_.astype(np.uint32) << np.uint32(1)) | signs
    is_def, is_true = random_assignments(rng, var_count, assignments, p_assigned)
    inst = Instance(var_count, lits, lens, is_def, is_true, assignments)
    if frac_false or frac_unit:
        plant(rng, inst, frac_false, frac_unit)
    return inst


def plain_64(seed: int = 1, scale: float = 1.0) -> Instance:
    """10k vars, 100k clauses, 64 assignments (60% assigned), 1% + 1% planted."""
    return make_instance(seed, int(10_000 * scale), int(100_000 * scale), 0.15, 100, 64, 0.6,
                         0.01, 0.01)


def sweep_1024(seed: int = 2, scale: float = 1.0) -> Instance:
    """1M vars, 4M clauses, 1024 assignments (40% assigned), 0.1% planted."""
    return make_instance(seed, max(1000, int(1_000_000 * scale)), int(4_000_000 * scale), 0.1,
                         200, 1024, 0.4, 0.0005, 0.0005)


def pareto_lengths(rng: np.random.Generator, n: int, alpha: float = 1.5, xmin: int = 2,
                   cap: int = 500) -> np.ndarray:
    """Heavy-tailed clause lengths: floor of Pareto(alpha) from xmin, capped."""
    u = 1.0 - rng.random(n)
    return np.minimum(cap, np.floor(xmin * u ** (-1.0 / alpha))).astype(np.int32)


def pareto_clauses(rng: np.random.Generator, n: int, var_count: int,
                   cap: int = 500) -> tuple[np.ndarray, np.ndarray]:
    lens = pareto_lengths(rng, n, cap=cap)
    vars_ = distinct_vars(rng, lens, var_count)
    signs = rng.integers(0, 2, size=vars_.size, dtype=np.uint32)
    return (vars_.astype(np.uint32) << np.uint32(1)) | signs, lens


@dataclass
class Step:
    """One batched step: new clauses, table updates, slot masks, roll-forward."""
    lits: np.ndarray
    lens: np.ndarray
    ups: np.ndarray        # VAR_UPDATE records (var, word, is_def, is_true)
    valid: np.ndarray      # uint64 [words]
    last_pos: np.ndarray   # int8 [words], -1 = word unused this step

    @property
    def assignments(self) -> int:
        return int(sum(int(v).bit_count() for v in self.valid))


class SolverStreams:
    """batched-16: solver streams sending uneven batches into one table.

    A planted total assignment sigma (uniform) stands for the solution the
    solvers converge to: every assigned variable takes sigma's value with
    probability `agree`, and every clause literal is true under sigma with
    probability `lit_true` (learned clauses are implied by the formula, so
    most of them are satisfied near a solution).  Random clauses against
    random assignments would make nearly every short clause unit in nearly
    every 64-slot word, which no solver population produces.

    Stream s owns words [s * wps, (s + 1) * wps) of the table (64 slots per
    word).  After every step each word holds one assignment in all its slots
    (the roll-forward copies the word's last sent slot), its base.  In a step
    stream s sends n_s ~ U[assig_lo, assig_hi] assignments into its first n_s
    slots; each one is the word's base with `flips` uniform variables redrawn
    (assigned with probability p_assigned), so only the touched (var, word)
    pairs are sent as updates.  It also inserts Poisson(clause_mean) clauses
    of Pareto(1.5) lengths from 2 capped at 500, distinct uniform variables.
    """

    def __init__(self, seed: int = 3, var_count: int = 1_000_000, streams: int = 16,
                 words_per_stream: int = 4, p_assigned: float = 0.95, agree: float = 0.99,
                 lit_true: float = 0.9, flips: int = 64, clause_mean: float = 2800.0,
                 assig_lo: int = 16, assig_hi: int = 256):
        if assig_hi > 64 * words_per_stream:
            raise ValueError("a stream sends more assignments than it has slots")
        self.rng = np.random.default_rng(seed)
        self.V = var_count
        self.streams = streams
        self.wps = words_per_stream
        self.words = streams * words_per_stream
        self.p_assigned = p_assigned
        self.agree = agree
        self.lit_true = lit_true
        self.flips = flips
        self.clause_mean = clause_mean
        self.assig_lo, self.assig_hi = assig_lo, assig_hi
        self.sigma = self.rng.random(var_count) < 0.5
        shape = (var_count, self.words)
        self.base_def = self.rng.random(shape, dtype=np.float32) < p_assigned
        agrees = self.rng.random(shape, dtype=np.float32) < agree
        self.base_true = self.base_def & (agrees == self.sigma[:, None])

    def initial_table(self) -> tuple[np.ndarray, np.ndarray]:
        ones = np.uint64(0xFFFFFFFFFFFFFFFF)
        d = np.where(self.base_def, ones, np.uint64(0))
        t = np.where(self.base_true, ones, np.uint64(0))
        return d, t

    def clauses(self, n: int) -> tuple[np.ndarray, np.ndarray]:
        """n clauses whose literals are each true under sigma w.p. lit_true."""
        lens = pareto_lengths(self.rng, n)
        vars_ = distinct_vars(self.rng, lens, self.V)
        st = self.rng.random(vars_.size) < self.lit_true
        # literal 2v + 1 is the negation: true under sigma iff sigma(v) is false
        sign = (st != self.sigma[vars_]).astype(np.uint32)
        return (vars_.astype(np.uint32) << np.uint32(1)) | sign, lens

    def initial_clauses(self, n: int) -> tuple[np.ndarray, np.ndarray]:
        return self.clauses(n)

    def _word_updates(self, w: int, k: int):
        rng, V, f = self.rng, self.V, self.flips
        slots = np.repeat(np.arange(k, dtype=np.int64), f)
        vs = rng.integers(0, V, size=k * f, dtype=np.int64)
        nd = rng.random(k * f) < self.p_assigned
        nt = nd & ((rng.random(k * f) < self.agree) == self.sigma[vs])
        key = vs * 64 + slots
        # a (var, slot) drawn twice keeps its last draw
        uk, ridx = np.unique(key[::-1], return_index=True)
        sel = key.size - 1 - ridx
        v_u = uk // 64
        bit = np.uint64(1) << (uk % 64).astype(np.uint64)
        uv, start = np.unique(v_u, return_index=True)
        changed = np.bitwise_or.reduceat(bit, start)
        dbits = np.bitwise_or.reduceat(np.where(nd[sel], bit, np.uint64(0)), start)
        tbits = np.bitwise_or.reduceat(np.where(nt[sel], bit, np.uint64(0)), start)
        ones = np.uint64(0xFFFFFFFFFFFFFFFF)
        bd = np.where(self.base_def[uv, w], ones, np.uint64(0))
        bt = np.where(self.base_true[uv, w], ones, np.uint64(0))
        is_def = (bd & ~changed) | dbits
        is_true = ((bt & ~changed) | tbits) & is_def
        last = np.uint64(k - 1)
        self.base_def[uv, w] = ((is_def >> last) & np.uint64(1)).astype(bool)
        self.base_true[uv, w] = ((is_true >> last) & np.uint64(1)).astype(bool)
        return uv, is_def, is_true

    def step(self) -> Step:
        from .engine import VAR_UPDATE_DTYPE
        rng = self.rng
        valid = np.zeros(self.words, dtype=np.uint64)
        last_pos = np.full(self.words, -1, dtype=np.int8)
        parts = []
        for s in range(self.streams):
            n = int(rng.integers(self.assig_lo, self.assig_hi + 1))
            for j in range(self.wps):
                k = min(64, n - 64 * j)
                if k <= 0:
                    break
                w = s * self.wps + j
                valid[w] = np.uint64(0xFFFFFFFFFFFFFFFF) if k == 64 else np.uint64((1 << k) - 1)
                last_pos[w] = k - 1
                uv, d, t = self._word_updates(w, k)
                u = np.empty(uv.size, VAR_UPDATE_DTYPE)
                u["var"] = uv
                u["word"] = w
                u["is_def"] = d
                u["is_true"] = t
                parts.append(u)
        ups = np.concatenate(parts) if parts else np.zeros(0, VAR_UPDATE_DTYPE)
        n_cl = int(rng.poisson(self.clause_mean * self.streams))
        lits, lens = self.clauses(n_cl)
        return Step(lits, lens, ups, valid, last_pos)
