"""The batched, serving and residency forms on their own workloads, at their
stated scale, against the CPU oracle.

* batched-16: the first steps of the 16-stream workload over 1M variables
  (1M initial clauses of Pareto(1.5) lengths up to 500, 64-word table)
  through gss_step_submit; every step's records equal the plain sequence's
  (append, apply, sweep, roll forward on a second engine) and the oracle's
  on the table that sequence had at sweep time (GpuRunner.cu:134-177).
* residency: the drifting hot-set workload over 2M variables with a small
  clause budget, under both policies; after every reduction the kept DB
  holds every clause of length <= 2 (Clauses.cu:255) and a sweep of it
  equals the oracle's on the DB and table read back from the device.
"""
import numpy as np
import pytest

import _oracle as orc
from gpusharesat_amd import residency as rs
from gpusharesat_amd import workloads
from gpusharesat_amd.engine import ClauseMap, Engine

pytestmark = pytest.mark.gpu


def records(hits, lens, rows):
    ids = ClauseMap(lens, rows).ids(hits)
    assert (ids >= 0).all()
    out = np.empty(hits.size, dtype=orc.HIT_DTYPE)
    out["clause"], out["word"] = ids, hits["word"]
    out["false_mask"], out["unit_mask"] = hits["false_mask"], hits["unit_mask"]
    return out[np.lexsort((out["word"], out["clause"]))]


def assert_same(got, exp, what):
    assert got.size == exp.size, (what, got.size, exp.size)
    for f in ("clause", "word", "false_mask", "unit_mask"):
        assert np.array_equal(got[f], exp[f]), (what, f)


@pytest.mark.parametrize("screen, signs, nsteps", [(False, "planted", 6), (True, "planted", 6),
                                                  (True, "random", 3)])
def test_batched16_steps_match_plain_and_oracle(screen, signs, nsteps):
    """batched-16 itself: graph steps, each equal to the plain sequence and
    to the oracle on the table at sweep time; with GSS_STEP_SCREEN the steps
    screen clauses against the table's aggregate first (refreshed from each
    step's updates after the first).  Also with random clause signs and
    assigned values (40 % assigned), where short clauses report in nearly
    every word and the screening drops almost nothing."""
    gen = workloads.SolverStreams(seed=3, var_count=1_000_000,
                                  **({} if signs == "planted" else
                                     dict(p_assigned=0.4, agree=0.5, lit_true=0.5)))
    d0, t0 = gen.initial_table()
    lits, lens = gen.initial_clauses(1_000_000)
    print(f"batched-16: 1M vars, {lens.size} initial clauses ({lits.size} literals), "
          f"{gen.words} words", flush=True)
    with Engine(0) as a, Engine(0) as b:
        for e in (a, b):
            base_rows = e.append(lits, lens)
            e.reset_assignments(gen.V, gen.words)
            e.upload(d0, t0)
        del d0, t0
        all_lits, all_lens, rows = [lits], [lens], [base_rows]
        for k in range(nsteps):
            st = gen.step()
            print(f"step {k}: {st.lens.size} clauses, {st.ups.size} updates, "
                  f"{st.assignments} assignments", flush=True)
            got_rows = a.step_enqueue(st.lits, st.lens, st.ups, st.valid, st.last_pos,
                                      screen=screen)
            n, _ = a.step_wait()
            step_hits = a.step_hits()
            assert step_hits.size == n
            new_rows = b.append(st.lits, st.lens)
            assert np.array_equal(got_rows, new_rows)
            all_lits.append(st.lits)
            all_lens.append(st.lens)
            rows.append(new_rows)
            b.apply(st.ups)
            d, t = b.download()
            cl_lens = np.concatenate(all_lens)
            cl_rows = np.concatenate(rows)
            off = np.zeros(cl_lens.size + 1, np.int64)
            np.cumsum(cl_lens, out=off[1:])
            exp = orc.sweep(off, np.concatenate(all_lits), d, t, st.valid)
            del d, t
            b.sweep(st.valid)
            assert_same(records(b.hits(), cl_lens, cl_rows), exp, f"plain step {k}")
            assert_same(records(step_hits, cl_lens, cl_rows), exp, f"batched step {k}")
            assert exp.size > 0
            b.roll_forward(st.ups, st.last_pos)


@pytest.mark.parametrize("policy", [rs.POLICY_ACTIVITY, rs.POLICY_LEARNED])
def test_residency_2m_vars_kept_db_checked_exactly(policy):
    cfg = rs.SimConfig(var_count=2_000_000, arrivals=1_400_000, rounds=7, budget=500_000,
                       redraw_every=3)
    checked = []

    def after_reduce(e: Engine, rnd: int, gone: np.ndarray):
        print(f"policy {policy} round {rnd}: {gone.size} evicted, DB {e.totals()}", flush=True)
        assert gone.size > 0
        lits, lens = [], []
        for length in range(1, 201):
            if not e.bucket_count(length):
                continue
            rows = e.read_bucket(length)
            if rows.size:
                lits.append(rows.ravel())
                lens.append(np.full(rows.shape[0], length, np.int32))
        cl_lits, cl_lens = np.concatenate(lits), np.concatenate(lens)
        n_short = int((cl_lens <= 2).sum())
        d, t = e.download()
        valid = np.full(e.words, np.uint64(0xFFFFFFFFFFFFFFFF))
        off = np.zeros(cl_lens.size + 1, np.int64)
        np.cumsum(cl_lens, out=off[1:])
        exp = orc.sweep(off, cl_lits, d, t, valid)
        n = e.sweep(valid)
        # rows in bucket order are exactly the concatenation read above
        rows = np.concatenate([np.arange(x.size, dtype=np.uint32) for x in lens])
        assert_same(records(e.hits(), cl_lens, rows), exp, f"kept DB after round {rnd}")
        assert n == exp.size > 0
        checked.append((rnd, cl_lens.size, n_short))

    out = rs.simulate(cfg, policy, after_reduce=after_reduce)
    assert out.reductions >= 2 and len(checked) == out.reductions
    # every clause of length <= 2 stays: their count never falls
    shorts = [c[2] for c in checked]
    assert all(b >= a for a, b in zip(shorts, shorts[1:]))
