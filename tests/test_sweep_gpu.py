"""Parity of the HIP sweep (through the C ABI) with the CPU oracle.

Bit-exact: the set of (clause, word, false mask, unit mask) records must equal
the oracle's on the same seeded input.  Covers every table width, clause
lengths 1..2000, duplicate and complementary literals, partial and empty slot
masks, an empty DB, the perfTest golden counts, the 1024-assignment workload at
full size, and a hit buffer forced too small (re-run, nothing dropped).
"""
import os

import numpy as np
import pytest

import _oracle as orc
from gpusharesat_amd import workloads
from gpusharesat_amd.engine import (ClauseMap, Engine, VAR_UPDATE_DTYPE, TUNE_FUSE_LEN,
                                    TUNE_MAX_HIT_CHUNKS, TUNE_TASK_CLAUSES, TUNE_UNROLL,
                                    TUNE_WAVES_PER_CU)

pytestmark = pytest.mark.gpu

ALL = np.uint64(0xFFFFFFFFFFFFFFFF)


def gpu_records(e: Engine, lens, rows):
    hits = e.hits()
    ids = ClauseMap(lens, rows).ids(hits)
    assert (ids >= 0).all()
    out = np.empty(hits.size, dtype=orc.HIT_DTYPE)
    out["clause"] = ids
    out["word"] = hits["word"]
    out["false_mask"] = hits["false_mask"]
    out["unit_mask"] = hits["unit_mask"]
    return out[np.lexsort((out["word"], out["clause"]))]


def run_both(e, lits, lens, var_count, is_def, is_true, valid):
    print(f"case: {lens.size} clauses, {lits.size} literals, {var_count} vars, "
          f"{is_def.shape[1]} words", flush=True)
    rows = e.append(lits, lens)
    e.reset_assignments(var_count, is_def.shape[1])
    e.upload(is_def, is_true)
    n = e.sweep(valid)
    got = gpu_records(e, lens, rows)
    assert got.size == n
    off = np.zeros(lens.size + 1, np.int64)
    off[1:] = np.cumsum(lens)
    exp = orc.sweep(off, lits, is_def, is_true, valid)
    return got, exp


def assert_same(got, exp):
    assert got.size == exp.size, (got.size, exp.size)
    for f in ("clause", "word", "false_mask", "unit_mask"):
        if not np.array_equal(got[f], exp[f]):
            bad = np.nonzero(got[f] != exp[f])[0][:5]
            raise AssertionError(f"field {f} differs at {bad}: {got[bad]} vs {exp[bad]}")


def random_case(seed, var_count, nclauses, maxlen, words, p_def, p_true=0.5, dup=0.05):
    rng = np.random.default_rng(seed)
    lens = rng.integers(1, maxlen + 1, size=nclauses).astype(np.int32)
    lits = rng.integers(0, 2 * var_count, size=int(lens.sum())).astype(np.uint32)
    # sprinkle duplicate and complementary literals
    off = np.zeros(nclauses + 1, np.int64)
    off[1:] = np.cumsum(lens)
    for c in np.nonzero((rng.random(nclauses) < dup) & (lens >= 2))[0]:
        lits[off[c] + 1] = lits[off[c]] ^ np.uint32(rng.integers(0, 2))
    A = words * 64
    d = workloads.bernoulli_words(rng, var_count, A, p_def)
    t = workloads.bernoulli_words(rng, var_count, A, p_true) & d
    valid = rng.integers(0, 2**63, size=words, dtype=np.uint64) | np.uint64(1)
    valid[rng.random(words) < 0.2] = ALL
    if words > 1:
        valid[rng.integers(0, words)] = 0
    return lits, lens, d, t, valid


@pytest.mark.parametrize("words", [1, 2, 4, 8, 16, 32, 64])
def test_parity_every_width(engine_factory, words):
    lits, lens, d, t, valid = random_case(100 + words, 300, 3000, 40, words, p_def=0.85,
                                          p_true=0.3)
    got, exp = run_both(engine_factory(), lits, lens, 300, d, t, valid)
    assert exp.size > 100
    assert_same(got, exp)


def test_parity_plain_64(engine_factory):
    inst = workloads.plain_64(seed=1)
    got, exp = run_both(engine_factory(), inst.lits, inst.lens, inst.var_count, inst.is_def,
                        inst.is_true, inst.valid)
    assert_same(got, exp)
    rec = {(int(c), int(w)): (int(f), int(u)) for c, w, f, u in
           zip(got["clause"], got["word"], got["false_mask"], got["unit_mask"])}
    p = inst.planted
    for c, kind, a in zip(p["clauses"], p["kinds"], p["assignments"]):
        fm, um = rec[(int(c), int(a) // 64)]
        assert (um if kind else fm) >> (int(a) % 64) & 1


def test_parity_long_clauses(engine_factory):
    # Literals are chosen false under a base assignment that the low 32 slots of
    # every word follow, so long clauses stay alive to the end of both passes
    # there; a few undefined variables turn some of them unit.
    rng = np.random.default_rng(7)
    V, words = 2000, 16
    base = rng.integers(0, 2, size=V).astype(np.uint64)  # 1 = var true
    lens = np.concatenate([rng.integers(100, 600, size=300), np.array([1000, 2000, 1]),
                           rng.integers(1, 4, size=200)]).astype(np.int32)
    vars_ = rng.integers(0, V, size=int(lens.sum()))
    lits = (2 * vars_ + base[vars_]).astype(np.uint32)
    low = np.uint64(0xFFFFFFFF)
    d = np.full((V, words), ALL, dtype=np.uint64)
    t = rng.integers(0, 2**63, size=(V, words), dtype=np.uint64) & ~low
    t |= np.where(base[:, None] == 1, low, np.uint64(0))
    for v in rng.choice(V, size=40, replace=False):
        d[v, rng.integers(0, words)] &= ~(np.uint64(1) << np.uint64(rng.integers(0, 32)))
    got, exp = run_both(engine_factory(), lits, lens, V, d, t & d, np.full(words, ALL))
    assert exp.size > 0 and (exp["false_mask"] != 0).sum() > 100
    assert_same(got, exp)


def test_parity_maximum_variable_count(engine_factory):
    # the largest table: its sentinel rows end at the last 32-bit stored row
    # offset; clauses use the first and the last variables of the range
    V = (1 << 24) - 1
    rng = np.random.default_rng(31)
    n = 4000
    lens = rng.integers(1, 9, size=n).astype(np.int32)
    vars_ = rng.integers(0, V, size=int(lens.sum()))
    vars_[:8] = [V - 1, V - 2, 0, 1, V - 1, 0, V - 3, 2]
    lits = (2 * vars_ + rng.integers(0, 2, size=vars_.size)).astype(np.uint32)
    d = rng.integers(0, np.iinfo(np.uint64).max, size=(V, 1), dtype=np.uint64, endpoint=True)
    t = rng.integers(0, np.iinfo(np.uint64).max, size=(V, 1), dtype=np.uint64,
                     endpoint=True) & d
    got, exp = run_both(engine_factory(), lits, lens, V, d, t, np.array([ALL]))
    assert exp.size > 0
    assert_same(got, exp)


def test_parity_maximum_clause_length(engine_factory):
    # clauses of the longest accepted length stay alive through both passes in
    # the low 32 slots of each word (every literal false there)
    rng = np.random.default_rng(32)
    V, words = 70000, 2
    base = rng.integers(0, 2, size=V).astype(np.uint64)
    lens = np.array([65535, 1, 65535, 3], np.int32)
    vars_ = rng.integers(0, V, size=int(lens.sum()))
    lits = (2 * vars_ + base[vars_]).astype(np.uint32)
    low = np.uint64(0xFFFFFFFF)
    d = np.full((V, words), ALL, dtype=np.uint64)
    t = rng.integers(0, 2**63, size=(V, words), dtype=np.uint64) & ~low
    t |= np.where(base[:, None] == 1, low, np.uint64(0))
    # one undefined variable in slot 5 of word 1 makes the clauses holding it
    # unit there instead of falsified
    v0 = int(vars_[0])
    d[v0, 1] &= ~np.uint64(1 << 5)
    got, exp = run_both(engine_factory(), lits, lens, V, d, t & d, np.full(words, ALL))
    assert exp.size > 0
    assert_same(got, exp)


def test_second_sweep_holds_only_its_own_records(engine_factory):
    """A sweep's records replace the previous sweep's: the reference's
    cleared reporter holds only the new reports
    (rel-newtech/test/OtherTests.cu:142-177)."""
    e = engine_factory()
    lens = np.ones(4, np.int32)
    rows = e.append(np.array([0, 2, 4, 6], np.uint32), lens)  # {x0} .. {x3}
    e.reset_assignments(4, 1)
    for false_vars in ((0, 1), (2, 3), ()):
        is_def = np.full((4, 1), ALL)
        is_true = is_def.copy()
        is_true[list(false_vars), 0] = 0
        e.upload(is_def, is_true)
        e.sweep(np.full(1, ALL))
        got = gpu_records(e, lens, rows)
        assert sorted(got["clause"].tolist()) == list(false_vars)
        assert (got["false_mask"] == ALL).all() and (got["unit_mask"] == 0).all()


def test_empty_db_and_empty_masks(engine_factory):
    e = engine_factory()
    e.reset_assignments(10, 1)
    assert e.sweep([ALL]) == 0
    assert e.hits().size == 0
    lits = np.array([0, 2, 4], np.uint32)
    lens = np.array([1, 2], np.int32)
    e.append(lits, lens)
    # all undefined: {x0} unit, {x1, x2} not
    assert e.sweep([0]) == 0
    n = e.sweep([ALL])
    h = e.hits()
    assert n == 1 and h[0]["len"] == 1 and int(h[0]["unit_mask"]) == 0xFFFFFFFFFFFFFFFF


def test_perftest_golden_on_gpu(engine_factory):
    off, lits, d, t = orc.perftest_instance(2_000_000, 2000)
    lens = np.diff(off).astype(np.int32)
    e = engine_factory()
    rows = e.append(lits, lens)
    e.reset_assignments(500, d.shape[1])
    e.upload(d, t)
    valid = np.zeros(32, np.uint64)
    valid[:31] = ALL
    valid[31] = np.uint64((1 << (2000 - 31 * 64)) - 1)
    e.sweep(valid)
    hits = e.hits()
    per = np.zeros(2048, np.int64)
    for h in hits:
        m = int(h["false_mask"]) | int(h["unit_mask"])
        while m:
            b = (m & -m).bit_length() - 1
            m &= m - 1
            per[int(h["word"]) * 64 + b] += 1
    cum = np.cumsum(per)
    import json
    known = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                        "reference_known_answers.json")))["perftest"]
    assert cum[14] == known["reports_after_15_rounds"]
    assert cum[1999] == known["reports_after_2000_rounds"]
    assert ClauseMap(lens, rows).ids(hits).min() >= 0


@pytest.mark.parametrize("scale", [0.25, 1.0])
def test_parity_sweep_1024(engine_factory, scale):
    inst = workloads.sweep_1024(seed=2, scale=scale)
    got, exp = run_both(engine_factory(), inst.lits, inst.lens, inst.var_count, inst.is_def,
                        inst.is_true, inst.valid)
    assert_same(got, exp)


def test_hit_buffer_regrowth_drops_nothing(engine_factory):
    inst = workloads.plain_64(seed=4, scale=0.5)
    e = engine_factory()
    e.set_tuning(TUNE_MAX_HIT_CHUNKS, 1)
    got, exp = run_both(e, inst.lits, inst.lens, inst.var_count, inst.is_def, inst.is_true,
                        inst.valid)
    assert exp.size > 5000
    assert_same(got, exp)


def test_overflow_rerun_refused_after_the_table_changed(engine_factory):
    """A sweep whose hit buffer overflows is re-run by the wait only when
    neither the DB nor the table changed after it was queued; with an update
    queued behind it, the wait fails instead of checking the updated table.
    Without a change the re-run's records equal the oracle's."""
    from gpusharesat_amd._lib import GssError
    inst = workloads.plain_64(seed=9, scale=0.2)
    e = engine_factory()
    e.set_tuning(TUNE_MAX_HIT_CHUNKS, 1)
    rows = e.append(inst.lits, inst.lens)
    e.reset_assignments(inst.var_count, inst.words)
    e.upload(inst.is_def, inst.is_true)
    e.sweep_async(inst.valid)
    ups = np.zeros(1, VAR_UPDATE_DTYPE)
    ups["var"], ups["word"], ups["is_def"] = 0, 0, ALL
    e.apply(ups)  # queued behind the sweep: the table changes
    with pytest.raises(GssError, match="cannot be re-run"):
        e.sweep_wait()
    # the same sweep with nothing queued behind it re-runs and drops nothing
    d, t = e.download()
    e.sweep_async(inst.valid)
    n = e.sweep_wait()
    exp = orc.sweep(inst.offsets, inst.lits, d, t, inst.valid)
    assert n == exp.size > 2000
    assert_same(gpu_records(e, inst.lens, rows), exp)


def test_tuning_never_changes_results(engine_factory):
    inst = workloads.plain_64(seed=6, scale=0.3)
    e = engine_factory()
    rows = e.append(inst.lits, inst.lens)
    e.reset_assignments(inst.var_count, 1)
    e.upload(inst.is_def, inst.is_true)
    e.sweep(inst.valid)
    ref = gpu_records(e, inst.lens, rows)
    for key, val in ((TUNE_WAVES_PER_CU, 4), (TUNE_TASK_CLAUSES, 4), (TUNE_TASK_CLAUSES, 1000),
                     (TUNE_WAVES_PER_CU, 0), (TUNE_TASK_CLAUSES, 0), (TUNE_FUSE_LEN, 1),
                     (TUNE_FUSE_LEN, 1000), (TUNE_FUSE_LEN, -1), (TUNE_FUSE_LEN, 0)):
        e.set_tuning(key, val)
        e.sweep(inst.valid)
        assert_same(gpu_records(e, inst.lens, rows), ref)


@pytest.mark.parametrize("words,unroll", [(16, 1), (16, 2), (16, 8), (32, 2), (32, 4),
                                          (64, 1), (64, 2), (64, 4)])
@pytest.mark.parametrize("fuse", [-1, 1, 3, 1000])
def test_parity_unroll_and_fuse(engine_factory, words, unroll, fuse):
    lits, lens, d, t, valid = random_case(300 + unroll * 7 + fuse + words, 400, 4000, 30, words,
                                          p_def=0.9, p_true=0.25)
    e = engine_factory()
    e.set_tuning(TUNE_UNROLL, unroll)
    e.set_tuning(TUNE_FUSE_LEN, fuse)
    got, exp = run_both(e, lits, lens, 400, d, t, valid)
    assert_same(got, exp)


@pytest.mark.parametrize("words", [1, 2, 4, 8])
def test_parity_without_fused_schedule(engine_factory, words):
    # fuse length -1 runs the kernel variant built without the fused schedule
    lits, lens, d, t, valid = random_case(900 + words, 400, 4000, 30, words, p_def=0.9,
                                          p_true=0.25)
    e = engine_factory()
    e.set_tuning(TUNE_FUSE_LEN, -1)
    got, exp = run_both(e, lits, lens, 400, d, t, valid)
    assert_same(got, exp)


def test_incremental_appends_and_bucket_replace(engine_factory):
    inst = workloads.plain_64(seed=8, scale=0.2)
    e = engine_factory()
    off = inst.offsets
    n = inst.lens.size
    rows = np.empty(n, np.uint32)
    for a, b in ((0, 10), (10, 5000), (5000, n)):
        rows[a:b] = e.append(inst.lits[off[a]:off[b]], inst.lens[a:b])
    assert e.totals() == (n, inst.lits.size)
    e.reset_assignments(inst.var_count, 1)
    e.upload(inst.is_def, inst.is_true)
    e.sweep(inst.valid)
    got = gpu_records(e, inst.lens, rows)
    exp = orc.sweep(off, inst.lits, inst.is_def, inst.is_true, inst.valid)
    assert_same(got, exp)
    # bucket contents round-trip
    L = int(inst.lens[0])
    sel = np.nonzero(inst.lens == L)[0]
    back = e.read_bucket(L)
    for c in sel[:50]:
        assert np.array_equal(back[rows[c]], inst.lits[off[c]:off[c + 1]])
    # replacing a bucket with its first half removes the other half's hits
    keep = back[: back.shape[0] // 2]
    e.set_bucket(L, keep.ravel())
    e.sweep(inst.valid)
    got2 = gpu_records(e, inst.lens, rows)
    dropped = set(sel[rows[sel] >= keep.shape[0]].tolist())
    exp2 = exp[~np.isin(exp["clause"], list(dropped))]
    assert_same(got2, exp2)


def test_assignment_apply_and_roll_forward(engine_factory):
    rng = np.random.default_rng(11)
    V, W = 500, 4
    e = engine_factory()
    e.reset_assignments(V, W)
    d0 = workloads.bernoulli_words(rng, V, W * 64, 0.5)
    t0 = workloads.bernoulli_words(rng, V, W * 64, 0.5) & d0
    e.upload(d0, t0)
    d, t = e.download()
    assert np.array_equal(d, d0) and np.array_equal(t, t0)
    sel = rng.choice(V * W, size=600, replace=False)
    ups = np.zeros(sel.size, VAR_UPDATE_DTYPE)
    ups["var"] = sel // W
    ups["word"] = sel % W
    ups["is_def"] = rng.integers(0, 2**63, size=sel.size, dtype=np.uint64)
    ups["is_true"] = rng.integers(0, 2**63, size=sel.size, dtype=np.uint64)
    e.apply(ups)
    d1, t1 = d0.copy(), t0.copy()
    d1[ups["var"], ups["word"]] = ups["is_def"]
    t1[ups["var"], ups["word"]] = ups["is_true"] & ups["is_def"]
    d, t = e.download()
    assert np.array_equal(d, d1) and np.array_equal(t, t1)
    last = [5, -1, 63, 0]
    e.roll_forward(ups, last)
    for w, p in enumerate(last):
        if p < 0:
            continue
        m = ups["word"] == w
        vs = ups["var"][m]
        bd = (d1[vs, w] >> np.uint64(p)) & np.uint64(1)
        bt = (t1[vs, w] >> np.uint64(p)) & np.uint64(1)
        d1[vs, w] = np.where(bd == 1, ALL, np.uint64(0))
        t1[vs, w] = np.where(bt == 1, ALL, np.uint64(0))
    d, t = e.download()
    assert np.array_equal(d, d1) and np.array_equal(t, t1)
    # growing the variable count keeps rows and adds undefined ones
    e.grow_vars(V + 100)
    d, t = e.download()
    assert np.array_equal(d[:V], d1) and not d[V:].any()


def test_invalid_arguments_are_errors(engine_factory):
    from gpusharesat_amd._lib import GssError
    e = engine_factory()
    with pytest.raises(GssError):  # empty clause
        e.append(np.array([], np.uint32), np.array([0], np.int32))
    with pytest.raises(GssError):  # literal beyond the table's variables
        e.reset_assignments(4, 1)
        e.append(np.array([8, 9], np.uint32), np.array([2], np.int32))
        e.sweep([ALL])
    e.clear()
    from gpusharesat_amd import _lib
    with pytest.raises(GssError):
        _lib.call("gss_assig_reset", e.handle, 10, 3)  # not a power of two
    with pytest.raises(GssError):  # sentinel rows beyond 32-bit stored offsets
        _lib.call("gss_assig_reset", e.handle, 1 << 24, 1)
    e.reset_assignments(10, 1)
    ups = np.zeros(1, VAR_UPDATE_DTYPE)
    ups["var"] = 10
    with pytest.raises(GssError):
        e.apply(ups)
