"""Host-side checks of the synthetic workload generators (no GPU)."""
import numpy as np

from gpusharesat_amd import workloads

ONES = np.uint64(0xFFFFFFFFFFFFFFFF)


def test_pareto_lengths_range_and_tail():
    lens = workloads.pareto_lengths(np.random.default_rng(0), 200_000)
    assert lens.min() == 2 and lens.max() == 500
    # P(L >= x) = (2 / x) ** 1.5 for the floor of Pareto(1.5) from 2
    for x in (4, 16, 64):
        assert abs((lens >= x).mean() - (2 / x) ** 1.5) < 0.01


def test_solver_streams_updates_and_roll_forward():
    """Applying a step's updates to the table and rolling every word forward
    to its last sent slot gives the generator's new bases; each sent slot is
    the previous base with at most `flips` variables redrawn."""
    g = workloads.SolverStreams(seed=5, var_count=3000, streams=4, words_per_stream=2,
                                flips=8, clause_mean=50, assig_lo=16, assig_hi=128)
    d, t = g.initial_table()
    for _ in range(4):
        prev_d, prev_t = d.copy(), t.copy()
        st = g.step()
        assert st.lens.size == 0 or st.lens.min() >= 2
        assert np.unique(st.ups["var"].astype(np.int64) * g.words + st.ups["word"]).size == \
            st.ups.size
        d[st.ups["var"], st.ups["word"]] = st.ups["is_def"]
        t[st.ups["var"], st.ups["word"]] = st.ups["is_true"]
        assert not (t & ~d).any()
        for w in range(g.words):
            k = int(st.last_pos[w]) + 1
            assert int(st.valid[w]) == (0 if k == 0 else (1 << k) - 1)
            for j in range(k):
                bd = (d[:, w] >> np.uint64(j)) & np.uint64(1)
                pd = (prev_d[:, w] >> np.uint64(j)) & np.uint64(1)
                bt = (t[:, w] >> np.uint64(j)) & np.uint64(1)
                pt = (prev_t[:, w] >> np.uint64(j)) & np.uint64(1)
                assert int(((bd != pd) | (bt != pt)).sum()) <= g.flips
            if k:
                p = np.uint64(k - 1)
                d[:, w] = np.where((d[:, w] >> p) & np.uint64(1), ONES, np.uint64(0))
                t[:, w] = np.where((t[:, w] >> p) & np.uint64(1), ONES, np.uint64(0))
        nd, nt = g.initial_table()
        assert np.array_equal(d, nd) and np.array_equal(t, nt)
