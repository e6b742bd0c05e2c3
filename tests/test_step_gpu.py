"""Batched steps (gss_step_*) through the C ABI.

A step appends clauses, applies assignment updates, sweeps every clause and
rolls the table forward, as one captured HIP graph per ping-pong slot that is
replayed with new data.  Each step must give exactly the records of the plain
sequence append -> apply -> sweep -> roll forward on a second engine, and the
oracle's records on the table that sequence had at sweep time; the table after
the steps must be the same on both engines.  Graphs are replayed many times,
with eager work on the same stream between the replays, with two steps in
flight, across bucket reallocation and across a table regrowth (which
re-captures the graphs).
"""
import numpy as np
import pytest

import _oracle as orc
from gpusharesat_amd import workloads
from gpusharesat_amd._lib import GssError
from gpusharesat_amd.engine import (ClauseMap, Engine, TUNE_FUSE_LEN, TUNE_MAX_HIT_CHUNKS,
                                    VAR_UPDATE_DTYPE)

pytestmark = pytest.mark.gpu

ALL = np.uint64(0xFFFFFFFFFFFFFFFF)


def make_step(rng, V, W, n_new, maxlen, n_ups, p_def=0.6, uniform=False):
    """uniform: every update sets all 64 slots of its word alike, so the
    slots of a word agree (the case the screening of batched steps prunes)."""
    lens = rng.integers(1, maxlen + 1, size=n_new).astype(np.int32)
    lits = rng.integers(0, 2 * V, size=int(lens.sum())).astype(np.uint32)
    sel = rng.choice(V * W, size=min(n_ups, V * W), replace=False)
    ups = np.zeros(sel.size, VAR_UPDATE_DTYPE)
    ups["var"] = sel // W
    ups["word"] = sel % W
    if uniform:
        ups["is_def"] = np.where(rng.random(sel.size) < p_def, ALL, np.uint64(0))
        ups["is_true"] = np.where(rng.random(sel.size) < 0.5, ALL, np.uint64(0))
    else:
        ups["is_def"] = workloads.bernoulli_words(rng, sel.size, 64, p_def)[:, 0]
        ups["is_true"] = (rng.integers(0, 2**63, size=sel.size, dtype=np.uint64)
                          << np.uint64(1))
    valid = rng.integers(0, 2**63, size=W, dtype=np.uint64) | np.uint64(1)
    valid[rng.random(W) < 0.3] = ALL
    if W > 1:
        valid[rng.integers(0, W)] = 0
    last_pos = rng.integers(-1, 64, size=W).astype(np.int8)
    return lits, lens, ups, valid, last_pos


def to_records(hits, lens_all, rows_all):
    ids = ClauseMap(lens_all, rows_all).ids(hits)
    assert (ids >= 0).all()
    out = np.empty(hits.size, dtype=orc.HIT_DTYPE)
    out["clause"] = ids
    out["word"] = hits["word"]
    out["false_mask"] = hits["false_mask"]
    out["unit_mask"] = hits["unit_mask"]
    return out[np.lexsort((out["word"], out["clause"]))]


def assert_same(got, exp, what):
    assert got.size == exp.size, (what, got.size, exp.size)
    for f in ("clause", "word", "false_mask", "unit_mask"):
        assert np.array_equal(got[f], exp[f]), (what, f)


class Plain:
    """The same steps through the plain entry points, checked by the oracle."""

    def __init__(self, e: Engine):
        self.e = e
        self.lits = np.zeros(0, np.uint32)
        self.lens = np.zeros(0, np.int32)
        self.rows = np.zeros(0, np.uint32)

    def step(self, lits, lens, ups, valid, last_pos):
        e = self.e
        if lens.size:
            r = e.append(lits, lens)
            self.lits = np.concatenate([self.lits, lits])
            self.lens = np.concatenate([self.lens, lens])
            self.rows = np.concatenate([self.rows, r])
        e.apply(ups)
        d, t = e.download()
        off = np.zeros(self.lens.size + 1, np.int64)
        off[1:] = np.cumsum(self.lens)
        exp = orc.sweep(off, self.lits, d, t, valid)
        n = e.sweep(valid)
        got = to_records(e.hits(), self.lens, self.rows)
        assert n == got.size
        assert_same(got, exp, "plain vs oracle")
        e.roll_forward(ups, last_pos)
        return exp


def run_steps(engine_factory, seed, V, W, nsteps, n_new, maxlen, n_ups, pipelined,
              grow_at=None, eager=True, tuning=(), page_limit=None, screen=False,
              uniform=False, read_in_flight=False, uniform_init=False):
    rng = np.random.default_rng(seed)
    a, b = engine_factory(), engine_factory()
    for key, val in tuning:
        a.set_tuning(key, val)
    if page_limit is not None:  # host copy endpoints pageable past this size
        a.set_page_locked_limit(page_limit)
        b.set_page_locked_limit(page_limit)
    a.reset_assignments(V, W)
    b.reset_assignments(V, W)
    if uniform_init:  # every word's 64 slots alike
        d0 = np.where(rng.random((V, W)) < 0.6, ALL, np.uint64(0))
        t0 = np.where(rng.random((V, W)) < 0.5, ALL, np.uint64(0)) & d0
    else:
        d0 = workloads.bernoulli_words(rng, V, W * 64, 0.5)
        t0 = workloads.bernoulli_words(rng, V, W * 64, 0.5) & d0
    a.upload(d0, t0)
    b.upload(d0, t0)
    plain = Plain(b)
    lens_a = np.zeros(0, np.int32)
    rows_a = np.zeros(0, np.uint32)
    expected = []
    waited = 0

    def check_one():
        nonlocal waited
        n, ms = a.step_wait()
        assert ms >= 0.0
        hits = a.step_hits()
        assert hits.size == n
        assert np.array_equal(a.step_hits_view(), hits)  # pinned view, same records
        got = to_records(hits, lens_a, rows_a)
        assert_same(got, expected[waited], f"step {waited} vs oracle")
        waited += 1

    for k in range(nsteps):
        if grow_at is not None and k == grow_at:
            while waited < k:
                check_one()
            V += 700
            a.grow_vars(V)
            b.grow_vars(V)
        nn = int(rng.integers(0, n_new + 1)) if k % 3 else n_new
        lits, lens, ups, valid, last_pos = make_step(rng, V, W, nn, maxlen, n_ups,
                                                     uniform=uniform)
        print(f"step {k}: {nn} clauses, {ups.size} updates, {V} vars, {W} words", flush=True)
        expected.append(plain.step(lits, lens, ups, valid, last_pos))
        rows = a.step_enqueue(lits, lens, ups, valid, last_pos,
                              screen=screen(k) if callable(screen) else screen)
        lens_a = np.concatenate([lens_a, lens])
        rows_a = np.concatenate([rows_a, rows])
        if not pipelined:
            check_one()
            if eager:  # eager work on the engine stream between replays
                da, ta = a.download()
                db, tb = b.download()
                assert np.array_equal(da, db) and np.array_equal(ta, tb), f"table after {k}"
        else:
            if read_in_flight:  # a plain call while steps are in flight
                da, ta = a.download()
                db, tb = b.download()
                assert np.array_equal(da, db) and np.array_equal(ta, tb), f"table at {k}"
            if waited < k:
                check_one()
    while waited < nsteps:
        check_one()
    # steps held to a few hit chunks overflow and re-run; the others (a
    # buffer of the worst case at these sizes) never do
    held = any(key == TUNE_MAX_HIT_CHUNKS and val for key, val in tuning)
    reruns = a.step_stats()["reruns"]
    assert reruns > 0 if held else reruns == 0, reruns
    da, ta = a.download()
    db, tb = b.download()
    assert np.array_equal(da, db) and np.array_equal(ta, tb)
    assert a.totals() == b.totals()
    for length in range(1, maxlen + 1):
        assert np.array_equal(a.read_bucket(length), b.read_bucket(length)), length


@pytest.mark.parametrize("W", [1, 4, 16, 64])
def test_steps_match_plain_and_oracle(engine_factory, W):
    run_steps(engine_factory, seed=300 + W, V=2000, W=W, nsteps=8, n_new=600, maxlen=30,
              n_ups=900, pipelined=False)


@pytest.mark.parametrize("uniform", [False, True])
@pytest.mark.parametrize("W", [1, 4, 16, 64])
def test_screened_steps_match_plain_and_oracle(engine_factory, W, uniform):
    """GSS_STEP_SCREEN: the same records as the plain sequence and the
    oracle, both where the screening keeps nearly every clause (slots drawn
    independently) and where it drops most (slots of a word alike), pipelined
    and across steps whose slot masks change."""
    run_steps(engine_factory, seed=1500 + W + 7 * uniform, V=2000, W=W, nsteps=8, n_new=600,
              maxlen=30, n_ups=900 if not uniform else 2000 * W, pipelined=True, screen=True,
              uniform=uniform)


@pytest.mark.parametrize("W", [8, 64])
def test_steps_without_fused_schedule(engine_factory, W):
    # the captured graph runs the sweep variant built without the fused schedule
    run_steps(engine_factory, seed=350 + W, V=2000, W=W, nsteps=5, n_new=600, maxlen=30,
              n_ups=900, pipelined=True, tuning=((TUNE_FUSE_LEN, -1),))


@pytest.mark.parametrize("W", [2, 16])
def test_two_steps_in_flight(engine_factory, W):
    run_steps(engine_factory, seed=400 + W, V=1500, W=W, nsteps=9, n_new=500, maxlen=20,
              n_ups=700, pipelined=True)


@pytest.mark.parametrize("W", [4, 64])
def test_records_read_after_the_slot_is_enqueued_again(engine_factory, W):
    # wait(k) -> enqueue(k + 2) into the same slot -> read step k's records:
    # they travel while the next step is staged and must not be overwritten
    rng = np.random.default_rng(450 + W)
    V = 1500
    a, b = engine_factory(), engine_factory()
    for e in (a, b):
        e.reset_assignments(V, W)
    d0 = workloads.bernoulli_words(rng, V, W * 64, 0.5)
    t0 = workloads.bernoulli_words(rng, V, W * 64, 0.5) & d0
    a.upload(d0, t0)
    b.upload(d0, t0)
    plain = Plain(b)
    steps = [make_step(rng, V, W, int(rng.integers(100, 800)), 25, 700) for _ in range(9)]
    expected, lens_a, rows_a = [], np.zeros(0, np.int32), np.zeros(0, np.uint32)
    todo = list(range(len(steps)))
    inflight = []

    def enqueue():
        nonlocal lens_a, rows_a
        k = todo.pop(0)
        lits, lens, ups, valid, last_pos = steps[k]
        expected.append(plain.step(lits, lens, ups, valid, last_pos))
        rows = a.step_enqueue(lits, lens, ups, valid, last_pos)
        lens_a = np.concatenate([lens_a, lens])
        rows_a = np.concatenate([rows_a, rows])
        inflight.append(k)

    enqueue()
    enqueue()
    while inflight:
        k = inflight.pop(0)
        n, _ = a.step_wait()
        if todo:
            enqueue()  # same slot as step k
        print(f"step {k}: {n} records, read after the next enqueue", flush=True)
        view = a.step_hits_view()
        assert view.size == n
        assert_same(to_records(view, lens_a, rows_a), expected[k], f"step {k}")
        assert_same(to_records(a.step_hits(), lens_a, rows_a), expected[k], f"step {k} copy")
    da, ta = a.download()
    db, tb = b.download()
    assert np.array_equal(da, db) and np.array_equal(ta, tb)


@pytest.mark.parametrize("pipelined", [False, True])
def test_steps_with_pageable_host_buffers(engine_factory, pipelined):
    # maxPageLockedMemory of 0 bytes: the records' host copies and the plain
    # path's staging are pageable (CorrespArr.cu:212-235); results unchanged
    run_steps(engine_factory, seed=380 + pipelined, V=1500, W=16, nsteps=6, n_new=500,
              maxlen=20, n_ups=600, pipelined=pipelined, page_limit=0)


def test_table_regrowth_recaptures(engine_factory):
    run_steps(engine_factory, seed=500, V=1200, W=8, nsteps=7, n_new=400, maxlen=12,
              n_ups=500, pipelined=True, grow_at=4)


@pytest.mark.parametrize("max_chunks", [0, 1])
def test_bucket_reallocation_mid_sequence(engine_factory, max_chunks):
    # steps of many short clauses make every bucket outgrow its capacity; with
    # one hit chunk per step, a step's re-run finds its rows moved by the
    # next step's growth
    run_steps(engine_factory, seed=600, V=800, W=4, nsteps=5, n_new=20000, maxlen=4,
              n_ups=300, pipelined=True,
              tuning=((TUNE_MAX_HIT_CHUNKS, max_chunks),) if max_chunks else ())


@pytest.mark.parametrize("read_in_flight", [False, True])
@pytest.mark.parametrize("W", [4, 64])
def test_overflowing_steps_are_re_run(engine_factory, W, read_in_flight):
    """Steps held to one hit chunk (GSS_TUNE_MAX_HIT_CHUNKS) overflow their
    buffers every time: the step's graph leaves the table as its sweep read it
    and halts the engine, so the step in flight behind it stages nothing; the
    wait (or a plain call made while steps are in flight) grows the buffer,
    re-runs the sweep and the roll forward, and replays the later step.
    Records and tables as the plain sequence and the oracle give them."""
    run_steps(engine_factory, seed=1700 + W + read_in_flight, V=2000, W=W, nsteps=8,
              n_new=600, maxlen=30, n_ups=900, pipelined=True,
              tuning=((TUNE_MAX_HIT_CHUNKS, 1),), read_in_flight=read_in_flight)


@pytest.mark.parametrize("pipelined", [False, True])
def test_screened_steps_keep_the_aggregate_current(engine_factory, pipelined):
    """Consecutive screened steps refresh the aggregate only where their
    updates changed the table (after the updates, and again after the roll
    forward); a step that is not screened, a plain read between steps or a
    table regrowth in between makes the next screened step rebuild it.  The
    slots of a word are alike, so most entries are set and a stale one would
    drop records."""
    run_steps(engine_factory, seed=1800 + pipelined, V=1500, W=16, nsteps=12, n_new=500,
              maxlen=25, n_ups=900, pipelined=pipelined, screen=lambda k: k % 5 != 3,
              uniform=True, uniform_init=True, grow_at=8)


def test_overflowing_screened_steps_are_re_run(engine_factory):
    # the replayed step rebuilds the aggregate the halted one left alone
    run_steps(engine_factory, seed=1750, V=2000, W=16, nsteps=6, n_new=600, maxlen=30,
              n_ups=2000 * 16, pipelined=True, screen=True, uniform=True,
              tuning=((TUNE_MAX_HIT_CHUNKS, 1),))


def test_sweep_only_step_and_errors(engine_factory):
    rng = np.random.default_rng(700)
    V, W = 400, 2
    a, b = engine_factory(), engine_factory()
    lits, lens, _, valid, _ = make_step(rng, V, W, 300, 9, 0)
    for e in (a, b):
        e.reset_assignments(V, W)
        d = workloads.bernoulli_words(np.random.default_rng(1), V, W * 64, 0.7)
        e.upload(d, d & workloads.bernoulli_words(np.random.default_rng(2), V, W * 64, 0.5))
    rows_a = a.append(lits, lens)
    rows_b = b.append(lits, lens)
    for _ in range(3):  # replays with nothing new: a sweep, table unchanged
        a.step_enqueue(valid=valid)
        n, _ = a.step_wait()
        got = to_records(a.step_hits(), lens, rows_a)
        b.sweep(valid)
        exp = to_records(b.hits(), lens, rows_b)
        assert n == got.size
        assert_same(got, exp, "sweep-only step")
    with pytest.raises(GssError):
        a.step_wait()  # nothing in flight
    with pytest.raises(GssError):  # literal without a table row
        a.step_enqueue(np.array([2 * V], np.uint32), np.array([1], np.int32), valid=valid)
    a.step_enqueue(valid=valid)
    a.step_enqueue(valid=valid)
    with pytest.raises(GssError):  # both slots busy
        a.step_enqueue(valid=valid)
    a.step_wait()
    a.step_wait()
    assert a.totals() == b.totals()


@pytest.mark.parametrize("threads", ["1", "3", "4"])
def test_large_steps_split_over_stage_threads(engine_factory, monkeypatch, threads):
    """Steps staging more than 1 MB (50k updates, up to 20k clauses) split
    their copies and clause records over the engine's stage threads
    (GPUSHARESAT_STAGE_THREADS, read when the engine is created); the records,
    rows and table must be those of the plain sequence and the oracle for any
    thread count, uneven part sizes included."""
    monkeypatch.setenv("GPUSHARESAT_STAGE_THREADS", threads)
    run_steps(engine_factory, seed=1300 + int(threads), V=20000, W=16, nsteps=4,
              n_new=20000, maxlen=30, n_ups=50000, pipelined=True)


def test_large_step_errors_leave_the_engine_unchanged(engine_factory, monkeypatch):
    """A large step whose last update or last literal is out of range is
    refused while its copies run on the stage threads; nothing of it reaches
    the engine, and the next step matches the plain sequence."""
    monkeypatch.setenv("GPUSHARESAT_STAGE_THREADS", "4")
    rng = np.random.default_rng(1400)
    V, W = 20000, 16
    a, b = engine_factory(), engine_factory()
    for e in (a, b):
        e.reset_assignments(V, W)
    d0 = workloads.bernoulli_words(rng, V, W * 64, 0.5)
    t0 = workloads.bernoulli_words(rng, V, W * 64, 0.5) & d0
    a.upload(d0, t0)
    b.upload(d0, t0)
    lits, lens, ups, valid, last_pos = make_step(rng, V, W, 20000, 30, 50000)
    before = a.totals()
    bad_ups = ups.copy()
    bad_ups["var"][-1] = V
    with pytest.raises(GssError, match="out of range"):
        a.step_enqueue(lits, lens, bad_ups, valid, last_pos)
    bad_lits = lits.copy()
    bad_lits[-1] = 2 * V
    with pytest.raises(GssError, match="grow the table"):
        a.step_enqueue(bad_lits, lens, ups, valid, last_pos)
    assert a.totals() == before
    assert np.array_equal(a.download()[0], d0)
    plain = Plain(b)
    exp = plain.step(lits, lens, ups, valid, last_pos)
    rows = a.step_enqueue(lits, lens, ups, valid, last_pos)
    n, _ = a.step_wait()
    got = to_records(a.step_hits(), lens, rows)
    assert n == got.size
    assert_same(got, exp, "large step after refused ones")
    da, ta = a.download()
    db, tb = b.download()
    assert np.array_equal(da, db) and np.array_equal(ta, tb)


def test_id_records_are_the_full_records_first_words(engine_factory):
    """A step submitted with GSS_RECORDS_IDS records, for every (clause, word)
    with a report, exactly the first 8-byte word of the gss_hit the full form
    records (row, length, word): the ids of a step equal the full step's
    records without their masks, on the same DB and table, for steps that
    append, update and roll forward in between; each form is refused by the
    other form's reader."""
    from gpusharesat_amd import serving as sv
    from gpusharesat_amd.engine import HIT_ID_DTYPE
    rng = np.random.default_rng(720)
    V, W = 1500, 16
    a, b = engine_factory(), engine_factory()
    for e in (a, b):
        e.reset_assignments(V, W)
    d0 = workloads.bernoulli_words(rng, V, W * 64, 0.5)
    t0 = workloads.bernoulli_words(rng, V, W * 64, 0.5) & d0
    a.upload(d0, t0)
    b.upload(d0, t0)
    plain = Plain(b)
    lens_a, rows_a = np.zeros(0, np.int32), np.zeros(0, np.uint32)
    for k in range(5):
        lits, lens, ups, valid, last_pos = make_step(rng, V, W, int(rng.integers(200, 900)), 25,
                                                     600)
        exp = plain.step(lits, lens, ups, valid, last_pos)
        rows = np.empty(lens.size, np.uint32)
        sv.submit(a, valid, lits=lits, lens=lens, ups=ups, last_pos=last_pos,
                  records=sv.RECORDS_IDS)
        n, _ = a.step_wait()
        ids = a.step_ids_view().copy()
        with pytest.raises(GssError, match="ids or clauses only"):
            a.step_hits_view()
        print(f"step {k}: {n} id records", flush=True)
        # the rows the step gave the new clauses: read back from the DB order
        lens_a = np.concatenate([lens_a, lens])
        rows_a = np.concatenate([rows_a, plain.rows[-lens.size:]])
        assert ids.size == n == exp.size
        # expected ids from the oracle's records: (row, length, word)
        want = np.empty(exp.size, HIT_ID_DTYPE)
        want["row"] = rows_a[exp["clause"]]
        want["len"] = lens_a[exp["clause"]]
        want["word"] = exp["word"]
        assert np.array_equal(np.sort(ids.view(np.uint64)), np.sort(want.view(np.uint64)))
    # the same state through a full-record step: its first words are the ids
    sv.submit(a, valid, records=sv.RECORDS_IDS)
    a.step_wait()
    ids = np.sort(a.step_ids_view().view(np.uint64))
    a.step_enqueue(valid=valid)
    a.step_wait()
    full = a.step_hits_view()
    with pytest.raises(GssError, match="another form"):
        a.step_ids_view()
    assert np.array_equal(ids, np.sort(full.view(np.uint64).reshape(-1, 3)[:, 0]))


@pytest.mark.parametrize("W, max_chunks, screen", [
    (1, 0, False), (4, 0, False), (16, 0, False), (64, 0, False), (16, 1, False), (64, 1, False),
    (4, 0, True), (16, 0, True), (64, 0, True), (64, 1, True)])
def test_clause_records_are_the_reports_by_clause(engine_factory, W, max_chunks, screen):
    """A step submitted with GSS_RECORDS_CLAUSES records one gss_hit_clause
    per clause with a report in some word: its row and length, and exactly
    the words of the oracle's reports for that clause; for steps that append,
    update and roll forward in between, at every width (sub-lanes below 16
    words, several words per lane above), also when one hit chunk per step
    makes every step overflow and re-run, and for screened steps (whose
    slots here are drawn independently, so most clauses pass the screen);
    the other readers refuse it."""
    from gpusharesat_amd import serving as sv
    rng = np.random.default_rng(730 + W)
    V = 1500
    a, b = engine_factory(), engine_factory()
    if max_chunks:
        a.set_tuning(TUNE_MAX_HIT_CHUNKS, max_chunks)
    for e in (a, b):
        e.reset_assignments(V, W)
    d0 = workloads.bernoulli_words(rng, V, W * 64, 0.5)
    t0 = workloads.bernoulli_words(rng, V, W * 64, 0.5) & d0
    a.upload(d0, t0)
    b.upload(d0, t0)
    plain = Plain(b)
    lens_a, rows_a = np.zeros(0, np.int32), np.zeros(0, np.uint32)
    for k in range(5):
        lits, lens, ups, valid, last_pos = make_step(rng, V, W, int(rng.integers(200, 900)), 25,
                                                     600)
        exp = plain.step(lits, lens, ups, valid, last_pos)
        sv.submit(a, valid, lits=lits, lens=lens, ups=ups, last_pos=last_pos,
                  records=sv.RECORDS_CLAUSES, flags=sv.STEP_SCREEN if screen else 0)
        n, _ = a.step_wait()
        cls = a.step_clauses_view().copy()
        with pytest.raises(GssError, match="ids or clauses only"):
            a.step_hits_view()
        with pytest.raises(GssError, match="another form"):
            a.step_ids_view()
        lens_a = np.concatenate([lens_a, lens])
        rows_a = np.concatenate([rows_a, plain.rows[-lens.size:]])
        # expected: one record per clause of the oracle's reports, its words ORed
        uc, inv = np.unique(exp["clause"], return_inverse=True)
        words = np.zeros(uc.size, np.uint64)
        np.bitwise_or.at(words, inv, np.uint64(1) << exp["word"].astype(np.uint64))
        print(f"W {W} step {k}: {n} clause records, {exp.size} reports", flush=True)
        assert cls.size == n == uc.size
        assert (cls["pad"] == 0).all()
        key_got = cls["row"].astype(np.uint64) | (cls["len"].astype(np.uint64) << np.uint64(32))
        key_exp = (rows_a[uc].astype(np.uint64)
                   | (lens_a[uc].astype(np.uint64) << np.uint64(32)))
        og, oe = np.argsort(key_got), np.argsort(key_exp)
        assert np.array_equal(key_got[og], key_exp[oe])
        assert np.array_equal(cls["words"][og], words[oe])
    assert (a.step_stats()["reruns"] > 0) == bool(max_chunks)


def test_step_with_no_valid_slot_still_appends_and_updates(engine_factory):
    """A step whose slots are all invalid checks nothing (no records) but
    still appends its clauses, applies its updates and rolls forward: the
    next step's records equal the oracle's on the table it then holds."""
    rng = np.random.default_rng(730)
    V, W = 900, 4
    a, b = engine_factory(), engine_factory()
    for e in (a, b):
        e.reset_assignments(V, W)
    plain = Plain(b)
    lits, lens, ups, valid, last_pos = make_step(rng, V, W, 700, 12, 500)
    plain.step(lits, lens, ups, np.zeros(W, np.uint64), last_pos)
    rows = a.step_enqueue(lits, lens, ups, np.zeros(W, np.uint64), last_pos)
    n, _ = a.step_wait()
    assert n == 0
    da, ta = a.download()
    db, tb = b.download()
    assert np.array_equal(da, db) and np.array_equal(ta, tb)
    exp = plain.step(np.zeros(0, np.uint32), np.zeros(0, np.int32), ups[:0], valid,
                     np.full(W, -1, np.int8))
    a.step_enqueue(valid=valid)
    n, _ = a.step_wait()
    assert n == exp.size > 0
    assert_same(to_records(a.step_hits(), lens, rows), exp, "after the empty-slot step")
