"""Serving: the steps it is built on (gss_step_submit with base planes, word
loads and patches) and the native serving loop (include/gss_serve.h).

* Base updates, word loads and patches leave the table exactly as the
  explicit per-slot assignments say (read back with gss_assig_download).
* A request's records are the same alone and in a mix (other requests
  around it, at another word offset), and equal the oracle's on the
  request's explicit assignments (its solver's base plus its patches):
  through steps, and through the native loop's C ABI, where the requests
  of a burst are admitted together on the device.
* A short run of the serving loop completes every request.
* serving-32 itself (32 solvers over 2M variables, Poisson arrivals, clause
  growth) for two seconds: sampled requests' records equal the oracle's on
  the DB and base their step had.
"""
import numpy as np
import pytest

import _oracle as orc
from gpusharesat_amd import _lib, serving as sv
from gpusharesat_amd.engine import TUNE_MAX_HIT_CHUNKS, ClauseMap

pytestmark = pytest.mark.gpu

ALL = np.uint64(0xFFFFFFFFFFFFFFFF)


def random_bases(rng, nb, V, p=0.7):
    d = rng.random((nb, V)) < p
    t = d & (rng.random((nb, V)) < 0.5)
    return d, t


def base_updates(d, t):
    ups = []
    for b in range(d.shape[0]):
        u = np.zeros(d.shape[1], sv.BASE_UP_DTYPE)
        u["var"] = np.arange(d.shape[1])
        u["base"] = b
        u["is_def"] = d[b]
        u["is_true"] = t[b]
        ups.append(u)
    return np.concatenate(ups)


def explicit(bd, bt, patches, nwords):
    """[V, nwords] planes of the request: base broadcast, then patches."""
    V = bd.size
    d = np.where(bd[:, None], ALL, np.uint64(0)) * np.ones((1, nwords), np.uint64)
    t = np.where(bt[:, None], ALL, np.uint64(0)) * np.ones((1, nwords), np.uint64)
    for p in patches:
        v, w, m = int(p["var"]), int(p["word"]), p["mask"]
        d[v, w] = (d[v, w] & ~m) | (p["is_def"] & m)
        t[v, w] = (t[v, w] & ~m) | (p["is_true"] & p["is_def"] & m)
    return d, t & d


def request_valid(k):
    nw = (k + 63) // 64
    v = np.full(nw, ALL)
    last = k - 64 * (nw - 1)
    if last < 64:
        v[-1] = np.uint64((1 << last) - 1)
    return v


def test_loads_and_patches_set_the_table(engine_factory):
    rng = np.random.default_rng(1)
    V, W = 500, 8
    e = engine_factory()
    e.reset_assignments(V, W)
    sv.bases_reset(e, 3)
    bd, bt = random_bases(rng, 3, V)
    sv.submit(e, np.zeros(W, np.uint64), base_ups=base_updates(bd, bt))
    e.step_wait()
    p = sv.slot_patches(rng, V, 128, 5, 0.6)  # relative words 0, 1
    p = p[p["word"] < 2]
    loads = np.zeros(2, sv.LOAD_DTYPE)
    loads["word"] = [2, 5]
    loads["base"] = [1, 0]
    pw = p.copy()
    pw["word"] = np.where(p["word"] == 0, 2, 5)
    sv.submit(e, np.zeros(W, np.uint64), loads=loads, patches=pw)
    e.step_wait()
    d, t = e.download()
    d2, t2 = explicit(bd[1], bt[1], p[p["word"] == 0], 1)
    d5, t5 = explicit(bd[0], bt[0], p[p["word"] == 1], 2)
    assert np.array_equal(d[:, 2], d2[:, 0]) and np.array_equal(t[:, 2], t2[:, 0])
    assert np.array_equal(d[:, 5], d5[:, 1]) and np.array_equal(t[:, 5], t5[:, 1])
    for w in (0, 1, 3, 4, 6, 7):  # never loaded: still undefined
        assert not d[:, w].any() and not t[:, w].any()
    # a later drift of base 1 reaches word 2 only when it is loaded again
    u = np.zeros(3, sv.BASE_UP_DTYPE)
    u["var"], u["base"], u["is_def"], u["is_true"] = [0, 1, 2], 1, [1, 0, 1], [1, 0, 0]
    sv.submit(e, np.zeros(W, np.uint64), base_ups=u)
    e.step_wait()
    assert np.array_equal(e.download()[0][:, 2], d2[:, 0])
    bd[1, :3], bt[1, :3] = [True, False, True], [True, False, False]
    ld = np.zeros(1, sv.LOAD_DTYPE)
    ld["word"], ld["base"] = 2, 1
    sv.submit(e, np.zeros(W, np.uint64), loads=ld)
    e.step_wait()
    d, t = e.download()
    assert np.array_equal(d[:, 2], np.where(bd[1], ALL, np.uint64(0)))
    assert np.array_equal(t[:, 2], np.where(bt[1], ALL, np.uint64(0)))


def test_request_records_alone_equal_in_mix_and_oracle(engine_factory):
    rng = np.random.default_rng(2)
    V, W, N = 400, 16, 3000
    e = engine_factory()
    lens = rng.integers(1, 13, size=N).astype(np.int32)
    lits = rng.integers(0, 2 * V, size=int(lens.sum())).astype(np.uint32)
    rows = e.append(lits, lens)
    e.reset_assignments(V, W)
    sv.bases_reset(e, 3)
    bd, bt = random_bases(rng, 3, V)
    sv.submit(e, np.zeros(W, np.uint64), base_ups=base_updates(bd, bt))
    e.step_wait()

    def req(solver, k, w0):
        p = sv.slot_patches(rng, V, k, 6, 0.6)
        nw = (k + 63) // 64
        last = k - 64 * (nw - 1)
        p = p[p["word"] < nw]
        if last < 64:
            tail = p["word"] == nw - 1
            p["mask"][tail] &= np.uint64((1 << last) - 1)
            p = p[p["mask"] != 0]
        return dict(solver=solver, k=k, w0=w0, patches=p)

    def step(reqs):
        valid = np.zeros(W, np.uint64)
        loads, patches = [], []
        for r in reqs:
            nw = (r["k"] + 63) // 64
            ld = np.zeros(nw, sv.LOAD_DTYPE)
            ld["word"] = np.arange(r["w0"], r["w0"] + nw)
            ld["base"] = r["solver"]
            loads.append(ld)
            p = r["patches"].copy()
            p["word"] += r["w0"]
            patches.append(p)
            valid[r["w0"]:r["w0"] + nw] = request_valid(r["k"])
        sv.submit(e, valid, loads=np.concatenate(loads), patches=np.concatenate(patches))
        e.step_wait()
        return e.step_hits()

    R = req(1, 200, 0)
    alone = step([R])
    X = req(0, 100, 0)
    Y = req(2, 64, 6)
    R2 = dict(R, w0=2)
    mix = step([X, R2, Y])
    part = mix[(mix["word"] >= 2) & (mix["word"] < 6)].copy()
    part["word"] -= 2
    key = ["len", "row", "word"]
    assert part.size == alone.size > 0
    a, b = np.sort(alone, order=key), np.sort(part, order=key)
    assert np.array_equal(a, b)
    # the oracle on R's explicit assignments
    d, t = explicit(bd[1], bt[1], R["patches"], 4)
    off = np.zeros(N + 1, np.int64)
    off[1:] = np.cumsum(lens)
    exp = orc.sweep(off, lits, d, t, request_valid(200))
    ids = ClauseMap(lens, rows).ids(alone)
    got = np.empty(alone.size, orc.HIT_DTYPE)
    got["clause"], got["word"] = ids, alone["word"]
    got["false_mask"], got["unit_mask"] = alone["false_mask"], alone["unit_mask"]
    got = got[np.lexsort((got["word"], got["clause"]))]
    assert got.size == exp.size
    for f in ("clause", "word", "false_mask", "unit_mask"):
        assert np.array_equal(got[f], exp[f])


def test_short_serving_run_completes_every_request():
    cfg = sv.ServeConfig(var_count=20_000, solvers=4, rate=300.0, duration=2.0,
                         clause_rate=20_000.0, pool=4, clause_pool=50_000, drift=16, flips=8)
    srv = sv.Server(cfg, log=print)
    try:
        out = srv.run()
    finally:
        srv.close()
    print(out.requests, out.steps, np.percentile(out.latency_ms, [50, 99]), flush=True)
    assert out.requests > 300 and out.steps > 0 and out.checks > 0
    assert (out.latency_ms > 0).all()


def _explicit_request(bd, bt, patches, k):
    """Planes [V, nw] of a request: its solver's base in every slot, then its
    patches (words relative to the request)."""
    nw = (k + 63) // 64
    V = bd.size
    d = np.repeat(np.where(bd, ALL, np.uint64(0))[:, None], nw, axis=1)
    t = np.repeat(np.where(bt, ALL, np.uint64(0))[:, None], nw, axis=1)
    if patches.size:
        v = patches["var"].astype(np.int64)
        w = patches["word"].astype(np.int64)
        m = patches["mask"]
        d[v, w] = (d[v, w] & ~m) | (patches["is_def"] & m)
        t[v, w] = (t[v, w] & ~m) | (patches["is_true"] & patches["is_def"] & m)
    assert V == d.shape[0]
    return d, t & d


def _oracle_request(off, lits, n_clauses, d, t, k):
    return orc.sweep(off[:n_clauses + 1], lits[:off[n_clauses]], d, t, request_valid(k))


def _as_oracle(recs):
    out = np.empty(recs.size, orc.HIT_DTYPE)
    out["clause"], out["word"] = recs["clause"], recs["word"]
    out["false_mask"], out["unit_mask"] = recs["false_mask"], recs["unit_mask"]
    return out[np.lexsort((out["word"], out["clause"]))]


def _same(got, exp, what):
    assert got.size == exp.size, (what, got.size, exp.size)
    for f in ("clause", "word", "false_mask", "unit_mask"):
        assert np.array_equal(got[f], exp[f]), (what, f)


@pytest.mark.parametrize("prefilter, max_chunks, W", [("1", 0, 16), ("0", 0, 16), ("1", 1, 16),
                                                      ("1", 0, 8), ("0", 0, 64)])
def test_native_loop_request_alone_equals_in_mix_and_oracle(monkeypatch, prefilter, max_chunks, W):
    """Through the C ABI: request R alone, then copies of R inside a burst
    of other requests (admitted together on the device, at other word
    offsets): every copy's records equal R alone, and the oracle's on R's
    explicit assignments.  With and without the serving steps' aggregate
    prefilter (GPUSHARESAT_SERVE_PREFILTER, read when the server's engine is
    created); the patches make literals false in many words, so the
    prefilter's rows are cleared by patches as well as set from the bases.
    With one hit chunk per step every step with records overflows: its wait
    re-runs its sweep and replays the step launched behind it, which found
    the engine halted and admitted nothing.  A serving step writes its loaded
    words in whole lines of the table (the first L words of each row, L =
    the loaded words rounded up to 16, at most the table's words): with 8
    words a variable's two rows share one line, with 64 a burst's words span
    several lines of each row."""
    monkeypatch.setenv("GPUSHARESAT_SERVE_PREFILTER", prefilter)
    rng = np.random.default_rng(21)
    V, N = 500, 4000
    lens = rng.integers(1, 13, size=N).astype(np.int32)
    lits = rng.integers(0, 2 * V, size=int(lens.sum())).astype(np.uint32)
    off = np.zeros(N + 1, np.int64)
    np.cumsum(lens, out=off[1:])
    bd, bt = random_bases(rng, 3, V)
    with sv.GpuServer(V, 3, words=W) as srv:
        if max_chunks:  # before the first step
            _lib.call("gss_set_tuning", srv.engine_handle(), TUNE_MAX_HIT_CHUNKS, max_chunks)
        assert srv.add_clauses(lits, lens) == 0
        srv.update_base(base_updates(bd, bt))

        def req(k):
            p = sv.slot_patches(rng, V, k, 6, 0.6)
            return p[p["word"] < (k + 63) // 64]

        R = req(200)

        def run(batch):
            ids = [srv.submit(s, k, p) for s, k, p in batch]
            assert None not in ids
            got = {}
            while len(got) < len(ids):
                for rid in srv.poll(64, 2_000_000):
                    got[rid] = srv.result(rid)
                    srv.release(rid)
            return [got[i] for i in ids]

        (meta, alone), = run([(1, 200, R)])
        assert meta["first_word"] == 0 and meta["clauses"] == N and alone.size > 0
        d, t = _explicit_request(bd[1], bt[1], R, 200)
        _same(_as_oracle(alone), _oracle_request(off, lits, N, d, t, 200), "R alone vs oracle")
        burst = []
        for _ in range(6):
            burst.append((0, int(rng.integers(1, 300)), None))
            burst.append((1, 200, R))
            burst.append((2, int(rng.integers(1, 200)), None))
        burst = [(s_, k, p if p is not None else req(k)) for s_, k, p in burst]
        out = run(burst)
        steps = {}
        for (s_, k, _), (m, _) in zip(burst, out):
            steps.setdefault(m["step"], []).append(m)
        mixed = 0
        for (s_, k, p), (m, recs) in zip(burst, out):
            if p is R:
                _same(_as_oracle(recs), _as_oracle(alone), f"R at word {m['first_word']}")
                if m["first_word"] > 0 and len(steps[m["step"]]) > 1:
                    mixed += 1
        assert mixed >= 1, steps
        st = srv.stats()
        assert st["requests_done"] == 1 + len(burst) and st["requests_pending"] == 0
        # (the serving thread is idle: every request is done)
        reruns = srv.step_stats()["reruns"]
        assert (reruns > 0) == bool(max_chunks), reruns


def test_native_loop_one_word_table():
    """A one-word table: each step admits one request of at most 64
    assignments and loads its word variable by variable (the whole-line
    form needs an even word count); every request equals the oracle on its
    explicit assignments."""
    rng = np.random.default_rng(29)
    V, N = 300, 2000
    lens = rng.integers(1, 10, size=N).astype(np.int32)
    lits = rng.integers(0, 2 * V, size=int(lens.sum())).astype(np.uint32)
    off = np.zeros(N + 1, np.int64)
    np.cumsum(lens, out=off[1:])
    bd, bt = random_bases(rng, 3, V)
    with sv.GpuServer(V, 3, words=1) as srv:
        assert srv.add_clauses(lits, lens) == 0
        srv.update_base(base_updates(bd, bt))
        sent = []
        for i in range(8):
            s_, k = i % 3, int(rng.integers(1, 65))
            p = sv.slot_patches(rng, V, k, 6, 0.6)
            p = p[p["word"] < 1]
            p["mask"] &= ALL if k == 64 else np.uint64((1 << k) - 1)
            p = p[p["mask"] != 0]
            rid = srv.submit(s_, k, p)
            assert rid is not None
            sent.append((rid, s_, k, p))
        got = {}
        while len(got) < len(sent):
            for rid in srv.poll(64, 2_000_000):
                got[rid] = srv.result(rid)
                srv.release(rid)
        for rid, s_, k, p in sent:
            meta, recs = got[rid]
            assert meta["first_word"] == 0 and meta["clauses"] == N
            d, t = _explicit_request(bd[s_], bt[s_], p, k)
            _same(_as_oracle(recs), _oracle_request(off, lits, N, d, t, k), f"request {rid}")


@pytest.mark.parametrize("records", [sv.RECORDS_IDS, sv.RECORDS_FULL])
def test_result_held_across_many_steps_keeps_its_records(records):
    """A client holds a finished request's result unread while 24 more
    requests run through their own steps: its records, lent from the
    engine's pinned buffers, are copied before those buffers are reused, and
    read late they still equal the oracle's on its explicit assignments (as
    do the records of requests read at once)."""
    rng = np.random.default_rng(23 + records)
    V, W, N = 500, 16, 4000
    lens = rng.integers(1, 13, size=N).astype(np.int32)
    lits = rng.integers(0, 2 * V, size=int(lens.sum())).astype(np.uint32)
    off = np.zeros(N + 1, np.int64)
    np.cumsum(lens, out=off[1:])
    bd, bt = random_bases(rng, 3, V)
    full = records == sv.RECORDS_FULL

    def check(meta, recs, solver, k, patches, what):
        d, t = _explicit_request(bd[solver], bt[solver], patches, k)
        exp = _oracle_request(off, lits, N, d, t, k)
        got = _as_oracle(recs)
        if full:
            _same(got, exp, what)
        else:  # id records: clause and word only
            assert got.size == exp.size, (what, got.size, exp.size)
            assert np.array_equal(got["clause"], exp["clause"]), what
            assert np.array_equal(got["word"], exp["word"]), what

    with sv.GpuServer(V, 3, words=W, records=records) as srv:
        assert srv.add_clauses(lits, lens) == 0
        srv.update_base(base_updates(bd, bt))

        def req(k):
            p = sv.slot_patches(rng, V, k, 6, 0.6)
            return p[p["word"] < (k + 63) // 64]

        def finish(rid):
            while rid not in done:
                done.update(srv.poll(64, 2_000_000))

        done = set()
        held_k = 200
        held_p = req(held_k)
        held = srv.submit(1, held_k, held_p)
        finish(held)
        for i in range(24):  # one request at a time: a step each
            s_, k = i % 3, int(rng.integers(1, 400))
            p = req(k)
            rid = srv.submit(s_, k, p)
            finish(rid)
            meta, recs = srv.result(rid)
            srv.release(rid)
            check(meta, recs, s_, k, p, f"request {i}")
        meta, recs = srv.result(held)
        assert meta["clauses"] == N and recs.size > 0
        check(meta, recs, 1, held_k, held_p, "held request read 24 steps later")
        srv.release(held)
        assert srv.stats()["steps"] >= 25


@pytest.mark.parametrize("planted, duration, min_requests", [(True, 2.0, 3000),
                                                           (False, 1.0, 1500)])
def test_serving32_workload_sampled_requests_match_oracle(planted, duration, min_requests):
    """serving-32 through the native loop for two seconds (32 solvers over
    2M variables, Poisson 2000 requests/s of U[64, 1024] assignments, clause
    growth at 200k/s): at least 50 sampled requests, each checked against
    the oracle on its explicit assignments (its solver's base after the base
    updates its step had applied, plus its patches) and the DB its step had.
    Also for one second with random clause signs and assigned values (the
    clause lengths of the 1024-assignment sweep, short clauses reporting in
    nearly every word)."""
    cfg = sv.ServeConfig(duration=duration, records=sv.RECORDS_FULL, planted=planted)
    srv = sv.Server(cfg, log=print, keep_history=True)
    big = []

    def keep(r):
        if r.k <= 128:
            return True
        if 448 < r.k <= 512 and len(big) < 3:
            big.append(r.idx)
            return True
        return False

    try:
        out = srv.run(keep=keep)
    finally:
        srv.close()
    print(out.requests, out.steps, np.percentile(out.latency_ms, [50, 99]), len(out.kept),
          flush=True)
    assert out.requests > min_requests and len(out.kept) >= 50
    # the DB: clause id i is pool clause i (no budget refusals, no wrap)
    for first, lo, hi in srv.clause_ids:
        assert first == lo
    off, lits = srv.cl_off, srv.cl_lits
    # replay the base updates in submission order, requests sorted by that count
    hist = srv.base_history
    cum = np.cumsum([0] + [h.size for h in hist])
    bd = np.zeros((cfg.solvers, cfg.var_count), bool)
    bt = np.zeros((cfg.solvers, cfg.var_count), bool)
    applied = 0
    items = sorted(out.kept.values(), key=lambda x: x[1]["base_updates"])[:60]
    checked = 0
    for r, meta, recs in items:
        need = meta["base_updates"]
        while applied < len(hist) and cum[applied + 1] <= need:
            u = hist[applied]
            # a variable drawn twice in one batch keeps its last update
            key = u["base"].astype(np.int64) * cfg.var_count + u["var"]
            _, last = np.unique(key[::-1], return_index=True)
            u = u[u.size - 1 - last]
            bd[u["base"], u["var"]] = u["is_def"].astype(bool)
            bt[u["base"], u["var"]] = (u["is_def"] & u["is_true"]).astype(bool)
            applied += 1
        assert cum[applied] == need, "a step applied part of one update batch"
        d, t = _explicit_request(bd[r.solver], bt[r.solver], srv.request_patches(r), r.k)
        exp = _oracle_request(off, lits, meta["clauses"], d, t, r.k)
        print(f"request {meta['request']} (solver {r.solver}, k {r.k}, step {meta['step']}, "
              f"{meta['clauses']} clauses): {recs.size} records", flush=True)
        _same(_as_oracle(recs), exp, f"request {meta['request']}")
        checked += 1
    assert checked >= 50
