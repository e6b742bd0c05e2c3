"""The residency scorer on the GPU (include/gss_residency.h).

* MLP: the device loss, gradient and scores equal the float64 CPU reference
  (tests/_mlp_ref.py, itself pinned by finite differences) within FP32
  accumulation tolerance.
* Activity policy: per-clause activities equal a float32 restatement of the
  reference's rule (new clause = increment after its decay step, one bump
  per record), the reduction keeps exactly the clauses the reference's
  approximate-median rule keeps, and the kept DB is still checked bit-exactly
  against the oracle.
* Features read back equal what the clause's history implies.
"""
import numpy as np
import pytest

import _mlp_ref as ref
import _oracle as orc
from gpusharesat_amd import residency as rs
from gpusharesat_amd import workloads
from gpusharesat_amd.engine import ClauseMap

pytestmark = pytest.mark.gpu

ALL = np.uint64(0xFFFFFFFFFFFFFFFF)


def make_res(engine_factory, max_tags=64):
    e = engine_factory()
    e.enable_tags()
    return e, rs.Residency(e, max_tags, seed=3)


def test_mlp_gradient_and_scores_match_cpu(engine_factory):
    _, r = make_res(engine_factory)
    rng = np.random.default_rng(5)
    p = rng.normal(0, 0.25, rs.PARAMS).astype(np.float32)
    r.params = p
    assert np.array_equal(r.params, p)
    n = 3000
    x = rng.normal(0, 1, (n, 12)).astype(np.float32)
    y = (rng.random(n) < 0.3).astype(np.float32)
    wt = (rng.random(n) + 0.5).astype(np.float32)
    loss, g = r.grad(x, y, wt)
    loss_ref, g_ref = ref.loss_and_grad(p, x, y, wt)
    print(f"loss {loss} vs {loss_ref}, max |g| {np.abs(g_ref).max()}", flush=True)
    assert abs(loss - loss_ref) <= 1e-4 * max(1.0, abs(loss_ref))
    tol = 1e-4 * np.abs(g_ref).max() + 1e-3 * np.abs(g_ref)
    bad = np.flatnonzero(np.abs(g - g_ref) > tol)
    assert bad.size == 0, (bad[:10], g[bad[:10]], g_ref[bad[:10]])
    # unweighted call
    _, g1 = r.grad(x[:500], y[:500])
    _, g1_ref = ref.loss_and_grad(p, x[:500], y[:500])
    assert np.allclose(g1, g1_ref, atol=1e-4 * np.abs(g1_ref).max(), rtol=1e-3)
    s = r.score(x)
    assert np.allclose(s, ref.prob(p, x), atol=2e-6, rtol=1e-5)


def sweep_tags(e, valid):
    """Records of a sweep mapped to clause tags: {tag: records}."""
    e.sweep(valid)
    hits = e.hits()
    tag_of = {int(L): e.read_tags(int(L)) for L in np.unique(hits["len"])}
    counts = {}
    for h in hits:
        t = int(tag_of[int(h["len"])][int(h["row"])])
        counts[t] = counts.get(t, 0) + 1
    return hits, counts


def test_record_accounting_with_64_words(engine_factory):
    """Short clauses against 64 dense words: most clauses have a record in
    nearly every word, so the per-wave merge of a clause's adjacent records
    sees runs of up to 64.  Counts, word bits, bumps and hit pairs are exact."""
    rng = np.random.default_rng(21)
    V, W = 200, 64
    e, r = make_res(engine_factory, max_tags=4000)
    decay = np.float32(0.99999)
    e.reset_assignments(V, W)
    d = workloads.bernoulli_words(rng, V, W * 64, 0.9)
    t = workloads.bernoulli_words(rng, V, W * 64, 0.3) & d
    e.upload(d, t)
    valid = np.full(W, ALL)
    n = 2000
    lens = rng.integers(1, 4, size=n).astype(np.int32)
    lits = rng.integers(0, 2 * V, size=int(lens.sum())).astype(np.uint32)
    tags = rng.permutation(4000)[:n].astype(np.uint32)
    r.add(lits, lens, tags, 0)
    incr = np.float32(1.0)
    act = {}
    for tg in tags:
        incr = np.float32(incr / decay)
        act[int(tg)] = incr
    words_cum = {}
    for rnd in range(2):
        hits, counts = sweep_tags(e, valid)
        assert max(counts.values()) == W  # some runs span the whole wave
        tag_of = {int(L): e.read_tags(int(L)) for L in np.unique(hits["len"])}
        words = {}
        for h in hits:
            tg = int(tag_of[int(h["len"])][int(h["row"])])
            words[tg] = words.get(tg, 0) | (1 << int(h["word"]))
        pairs_exp = int(np.bitwise_count(hits["false_mask"] | hits["unit_mask"]).sum())
        assert r.account(rnd) == pairs_exp
        for tg, c in counts.items():
            a = act[tg]
            for _ in range(c):
                a = np.float32(a + incr)
            act[tg] = a
            words_cum[tg] = words_cum.get(tg, 0) | words[tg]
        tg_all = np.array(sorted(act), dtype=np.uint32)
        got = r.activity(tg_all)
        assert np.array_equal(got, np.array([act[int(x)] for x in tg_all], np.float32))
        f = r.features(rnd, tg_all)
        cnt = np.array([counts.get(int(x), 0) for x in tg_all])
        assert np.allclose(f[:, 4], np.log1p(cnt) / 6, atol=1e-6)
        pc = np.array([bin(words_cum.get(int(x), 0)).count("1") for x in tg_all])
        assert np.array_equal(f[:, 10], (pc / 64).astype(np.float32))


def test_activity_policy_and_kept_db_parity(engine_factory):
    rng = np.random.default_rng(9)
    V, W = 300, 2
    e, r = make_res(engine_factory, max_tags=5000)
    decay = np.float32(0.99999)
    e.reset_assignments(V, W)
    d = workloads.bernoulli_words(rng, V, W * 64, 0.85)
    t = workloads.bernoulli_words(rng, V, W * 64, 0.4) & d
    e.upload(d, t)
    valid = np.full(W, ALL)
    incr = np.float32(1.0)
    act = {}
    lens_of, lits_of = {}, {}
    tags_free = rng.permutation(5000).astype(np.uint32)
    used = 0
    for rnd in range(3):
        n = 700
        lens = rng.integers(1, 9, size=n).astype(np.int32)
        lits = rng.integers(0, 2 * V, size=int(lens.sum())).astype(np.uint32)
        tags = tags_free[used:used + n]
        used += n
        r.add(lits, lens, tags, rnd)
        off = np.concatenate([[0], np.cumsum(lens)])
        for i, tg in enumerate(tags):
            incr = np.float32(incr / decay)
            act[int(tg)] = incr
            lens_of[int(tg)] = int(lens[i])
            lits_of[int(tg)] = lits[off[i]:off[i + 1]]
        _, counts = sweep_tags(e, valid)
        pairs = r.account(rnd)
        for tg, c in counts.items():
            a = act[tg]
            for _ in range(c):
                a = np.float32(a + incr)
            act[tg] = a
        assert pairs > 0
        tg_all = np.array(sorted(act), dtype=np.uint32)
        got = r.activity(tg_all)
        exp = np.array([act[int(x)] for x in tg_all], dtype=np.float32)
        assert np.array_equal(got, exp), f"round {rnd}"
        f = r.features(rnd, tg_all)
        L = np.array([lens_of[int(x)] for x in tg_all])
        assert np.allclose(f[:, 0], np.log2(L) / 9, atol=1e-6)
        assert np.array_equal(f[:, 1], (L <= 2).astype(np.float32))
        hit_now = np.array([counts.get(int(x), 0) > 0 for x in tg_all])
        assert np.array_equal(f[:, 3] > 0.5, hit_now)
        cnt = np.array([counts.get(int(x), 0) for x in tg_all])
        assert np.allclose(f[:, 4], np.log1p(cnt) / 6, atol=1e-6)
    # reduction: the reference's approximate median over log(activity)
    tg_all = np.array(sorted(act), dtype=np.uint32)
    a = np.array([act[int(x)] for x in tg_all], dtype=np.float32)
    lowest = np.float32(np.log(np.finfo(np.float32).tiny))
    largest = np.float32(np.log(np.finfo(np.float32).max))
    step = np.float32((largest - lowest) / np.float32(20000))
    b = np.floor((np.log(a).astype(np.float64) - np.float64(lowest)) / np.float64(step))
    b = np.clip(b, 0, 19999).astype(np.int64)
    hist = np.bincount(b, minlength=20000)
    k = int(np.searchsorted(np.cumsum(hist), a.size // 2))
    thr = np.float32(np.exp(np.float32(lowest + np.float32(k + 1) * step)))
    L = np.array([lens_of[int(x)] for x in tg_all])
    keep_exp = (L <= 2) | (a >= thr)
    ev, kept = r.reduce(rs.POLICY_ACTIVITY, 3)
    assert np.array_equal(~ev[tg_all], keep_exp)
    assert kept == int(keep_exp.sum()) and 0 < kept < a.size
    # the kept DB: every bucket holds exactly the kept clauses, and a sweep
    # of it equals the oracle over them
    kept_tags = set(int(x) for x in tg_all[keep_exp])
    lits_k, lens_k, rows_k, ids_k = [], [], [], []
    for length in range(1, 9):
        rows = e.read_bucket(length)
        tg = e.read_tags(length)
        assert {int(x) for x in tg} == {x for x in kept_tags if lens_of[x] == length}
        for row, x in enumerate(tg):
            assert np.array_equal(rows[row], lits_of[int(x)])
            lits_k.append(lits_of[int(x)])
            lens_k.append(length)
            rows_k.append(row)
    lens_k = np.array(lens_k, np.int32)
    lits_k = np.concatenate(lits_k).astype(np.uint32)
    e.sweep(valid)
    hits = e.hits()
    ids = ClauseMap(lens_k, np.array(rows_k)).ids(hits)
    got = np.empty(hits.size, dtype=orc.HIT_DTYPE)
    got["clause"], got["word"] = ids, hits["word"]
    got["false_mask"], got["unit_mask"] = hits["false_mask"], hits["unit_mask"]
    got = got[np.lexsort((got["word"], got["clause"]))]
    off = np.concatenate([[0], np.cumsum(lens_k)]).astype(np.int64)
    exp = orc.sweep(off, lits_k, d, t, valid)
    assert got.size == exp.size
    for fld in ("clause", "word", "false_mask", "unit_mask"):
        assert np.array_equal(got[fld], exp[fld])


def test_online_training_learns(engine_factory):
    cfg = rs.SimConfig(var_count=20_000, hot_vars=600, arrivals=240_000, rounds=40,
                       budget=60_000, redraw_every=15, seed=21)
    out = rs.simulate(cfg, rs.POLICY_LEARNED, log=print)
    print(out, out.hits_per_gb, flush=True)
    assert out.reductions > 0 and out.hit_pairs > 0
    assert 0.0 < out.train_loss < 0.69  # below the loss of a constant predictor


def test_learned_reduction_keeps_short_clauses_and_half_of_the_rest(engine_factory):
    rng = np.random.default_rng(13)
    V, W, n = 500, 1, 4000
    e, r = make_res(engine_factory, max_tags=n)
    e.reset_assignments(V, W)
    e.upload(workloads.bernoulli_words(rng, V, 64, 0.5), np.zeros((V, W), np.uint64))
    lens = rng.integers(1, 9, size=n).astype(np.int32)
    lits = rng.integers(0, 2 * V, size=int(lens.sum())).astype(np.uint32)
    tags = np.arange(n, dtype=np.uint32)
    r.add(lits, lens, tags, 0)
    e.sweep(np.full(W, ALL))
    r.account(0)
    ev, kept = r.reduce(rs.POLICY_LEARNED, 0)
    assert not ev[lens <= 2].any()
    long_ = lens > 2
    frac = ev[long_].mean()
    assert 0.25 < frac < 0.75, frac
    assert kept == n - int(ev.sum()) == e.totals()[0]


def test_learned_policy_gains_over_activity_policy():
    """The trained network against the reference's activity median on the
    same seeded drifting-hot-set workload (scale 0.05: 100k variables, 3M
    arrivals, 1M-clause budget, 300 rounds): more hit pairs per GB checked."""
    cfg = rs.SimConfig(var_count=100_000, arrivals=3_000_000, budget=1_000_000, seed=11)
    act = rs.simulate(cfg, rs.POLICY_ACTIVITY)
    lrn = rs.simulate(cfg, rs.POLICY_LEARNED)
    print(f"activity {act.hits_per_gb:.4e}, learned {lrn.hits_per_gb:.4e} hit pairs/GB", flush=True)
    assert lrn.reductions > 0 and act.reductions > 0
    assert lrn.hits_per_gb >= 1.02 * act.hits_per_gb
