"""The bound on sweep-1024's roofline fraction, from the sweep's own schedule
replayed on the CPU (no GPU).

The 16-word sweep checks a clause by reading 128-byte table rows: clauses of
at most 4 literals read both rows of every literal in one fused pass; longer
ones read the not-false row of their literals 4 at a time (one step) until
every valid slot holds two non-false literals, then, only if some slot is
still open, the negated rows of pass 2.  Replaying that schedule on clauses
of sweep-1024's length distribution against 1024 random assignments (40 %
assigned) gives the rows a clause must read; times 128 bytes, plus the
literal stream, that is the sweep's traffic whatever the kernel does, since
concurrently checked clauses share no rows (their variables are uniform over
2M literal rows).  Against the algorithmic bytes (each table row once, the
literals, the records) and the device's random-line gather rate it bounds
the fraction of peak HBM bandwidth a clause-by-clause check can reach.
"""
import numpy as np

import _oracle as orc
from gpusharesat_amd import workloads

ALL = np.uint64(0xFFFFFFFFFFFFFFFF)


def rows_read(lits, lens, is_def, is_true, unroll=4, fuse_len=4):
    """Rows of 128 bytes the sweep's schedule reads for each clause."""
    nf_pos = ~(is_def & ~is_true)  # slots where literal +v is not false
    nf_neg = ~(is_def & is_true)   # slots where literal -v is not false
    out = np.empty(lens.size, np.int64)
    off = np.zeros(lens.size + 1, np.int64)
    np.cumsum(lens, out=off[1:])
    for c in range(lens.size):
        n = int(lens[c])
        if n <= fuse_len:
            out[c] = 2 * n  # both rows of each literal, one pass
            continue
        cl = lits[off[c]:off[c + 1]]
        v, neg = cl >> 1, (cl & 1).astype(bool)
        rows = np.where(neg[:, None], nf_neg[v], nf_pos[v])
        c1 = np.zeros(rows.shape[1], np.uint64)
        c2 = np.zeros(rows.shape[1], np.uint64)
        read = 0
        dead = False
        for s in range(0, n, unroll):
            for r in rows[s:s + unroll]:
                c2 |= c1 & r
                c1 |= r
            read = min(n, s + unroll)
            if np.all(c2 == ALL):
                dead = True
                break
        if not dead:
            read += n  # pass 2 (at most every negated row)
        out[c] = read
    return out


def test_rows_per_clause_bound_the_roofline_fraction():
    rng = np.random.default_rng(5)
    V, N, A = 20_000, 20_000, 1024
    lens = workloads.clause_lengths(rng, N, 0.1, 200)
    vars_ = workloads.distinct_vars(rng, lens, V)
    lits = (vars_.astype(np.uint32) << np.uint32(1)) | rng.integers(0, 2, size=vars_.size,
                                                                    dtype=np.uint32)
    is_def, is_true = workloads.random_assignments(rng, V, A, 0.4)
    rows = rows_read(lits, lens, is_def, is_true)
    per_clause = rows.mean()
    # sweep-1024 reads 8.0 rows per clause by this count, and the rows are
    # the fetch its PMC pass measures (4.24 GB for 4M clauses with the
    # literals); a finer vote reads fewer but waits on more steps
    assert 7.7 < per_clause < 8.3, per_clause
    off = np.zeros(N + 1, np.int64)
    np.cumsum(lens, out=off[1:])
    head = slice(0, int(off[2000]))
    assert rows_read(lits[head], lens[:2000], is_def, is_true, unroll=1).mean() <= \
        rows_read(lits[head], lens[:2000], is_def, is_true, unroll=4).mean()
    # records per clause from the oracle (at most one per (clause, word))
    recs = orc.sweep(off, lits, is_def, is_true, np.full(A // 64, ALL)).size / N
    mean_len = lens.mean()
    # per clause at sweep-1024's proportions: 1M variables for 4M clauses,
    # each variable's two 128-byte rows once
    algorithmic = 4 * mean_len + 256 * (1_000_000 / 4_000_000) + 24 * recs
    traffic = 128 * per_clause + 4 * mean_len
    gather_ceiling, peak = 7.0e12, 8.0e12  # random 128-byte lines, 256 MB table
    bound = algorithmic / traffic * gather_ceiling / peak
    print(f"{per_clause:.2f} rows per clause, {recs:.2f} records per clause: "
          f"algorithmic {algorithmic:.0f} B, traffic {traffic:.0f} B per clause, "
          f"roofline fraction at most {bound:.3f}", flush=True)
    assert 0.15 < bound < 0.19, bound
