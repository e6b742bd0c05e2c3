"""GpuClauseSharer API on the GPU: the reference's library-level tests restated
through the Python mirror of include/gpuShareLib/GpuClauseSharer.h, plus a
randomised end-to-end check of the assignment snapshot semantics against the
CPU oracle.

Reference tests (glucose-syrup/test/GpuSolverTest.cu unless noted):
  testReported :285-328, testClausesAssigsReported :343-391,
  testFindClausesMultiThread :394-454, testOneAssignmentThenTwo :486-508,
  SolverDoesntImportSameClauseTwice(+OnSuccessiveGpuExecutions) :511-563,
  SolverHasManyClausesReportedAllAtOnce :678-724, testReduceDb :889-935,
  testSolverPassesManyAssignments :949-981, testManyAssignments :157-181,
  SolverCanReimportClause (rel-newtech/test/GpuSolverTest.cu:577-596),
  TwoSolverImportBinary :605-632, SolverUnsets :634-654,
  SolverImportUnary :656-676, SolverHasManyClausesReported :678-694,
  OneInstanceTwoAssignments :727-765, findConflict :858-887 and
  testSendClauseToGpu :999-1035 (the parts visible through the sharer),
  the threading contract of GpuClauseSharer.h:78-158 (solver threads and
  the GPU thread at once),
  ClauseActivityLbdTest.cu:59-95 (activity init / bump / decay).
"""
import numpy as np
import pytest

import _oracle as orc
from gpusharesat_amd.sharer import GpuClauseSharer, default_options, mkLit

pytestmark = pytest.mark.gpu


def sharer(var_count, solvers, **opts):
    s = GpuClauseSharer(default_options(min_gpu_latency_micros=0, verbosity=0, **opts))
    s.setVarCount(var_count)
    s.setCpuSolverCount(solvers)
    return s


def T(v):
    return mkLit(v)


def F(v):
    return mkLit(v, True)


def test_reported_unit_clause():
    with sharer(10, 3) as s:
        cid = s.addClause(-1, [mkLit(3)])
        assert s.trySetSolverValues(0, [F(3)])
        assert s.trySendAssignment(0) == 0
        s.execute()
        assert s.popReportedClause(0) == ([mkLit(3)], cid)
        assert s.popReportedClause(0) is None
        assert cid == 0


def test_clauses_assigs_reported():
    with sharer(3, 3) as s:
        for v in range(3):
            s.addClause(-1, [mkLit(v)])
        assert s.trySetSolverValues(0, [F(0), T(1), T(2)])
        s.trySendAssignment(0)
        assert s.trySetSolverValues(0, [T(0), F(1), F(2)])
        s.trySendAssignment(0)
        assert s.trySetSolverValues(1, [T(0), F(1), T(2)])
        s.trySendAssignment(1)
        s.execute()
        assert [len(s.popAll(i)) for i in range(3)] == [3, 1, 0]
        assert s.trySetSolverValues(0, [T(1)])
        s.trySendAssignment(0)
        s.execute()
        assert len(s.popAll(0)) == 1
        assert len(s.popAll(1)) == 0


def test_find_clauses():
    with sharer(3, 1) as s:
        assert s.trySetSolverValues(0, [F(0), T(1)])  # var 2 stays undefined
        s.trySendAssignment(0)
        s.addClause(-1, [mkLit(0), mkLit(1)])
        s.addClause(-1, [mkLit(0), F(1)])
        s.addClause(-1, [F(1), mkLit(2)])
        s.execute()
        got = sorted(tuple(l) for l, _ in s.popAll(0))
        assert got == sorted([(mkLit(0), F(1)), (F(1), mkLit(2))])


def test_one_assignment_then_two():
    with sharer(4, 1) as s:
        s.addClause(-1, [F(0), F(1), mkLit(2)])
        assert s.trySetSolverValues(0, [T(0)])
        s.trySendAssignment(0)
        s.execute()
        assert s.trySetSolverValues(0, [T(3)])
        s.trySendAssignment(0)
        assert s.trySetSolverValues(0, [T(1)])
        s.trySendAssignment(0)
        s.execute()
        assert len(s.popAll(0)) == 1


def test_same_clause_imported_once():
    with sharer(3, 1) as s:
        assert s.trySetSolverValues(0, [T(0)])
        s.trySendAssignment(0)
        s.unsetSolverValues(0, [T(0)])
        assert s.trySetSolverValues(0, [F(1)])
        s.trySendAssignment(0)
        s.addClause(-1, [F(0), mkLit(1)])
        s.execute()
        assert len(s.popAll(0)) == 1
        assert s.getOneSolverStat(0, "reportedClauses") == 1


def test_same_clause_imported_once_across_runs():
    with sharer(3, 1) as s:
        s.addClause(-1, [F(0), mkLit(1)])
        assert s.trySetSolverValues(0, [T(0)])
        s.trySendAssignment(0)
        s.gpuRun()
        s.unsetSolverValues(0, [T(0)])
        assert s.trySetSolverValues(0, [F(1)])
        s.trySendAssignment(0)
        s.gpuRun()
        n = len(s.popAll(0))
        s.gpuRun()
        n += len(s.popAll(0))
        assert n == 1
        assert s.getOneSolverStat(0, "reportedClauses") == 1


def test_same_assignment_resent_reports_again():
    with sharer(5, 1) as s:
        s.addClause(-1, [F(0), mkLit(1)])
        assert s.trySetSolverValues(0, [T(0)])
        assert s.trySendAssignment(0) >= 0
        s.execute()
        assert len(s.popAll(0)) == 1
        assert s.trySendAssignment(0) >= 0
        s.execute()
        assert len(s.popAll(0)) == 1


def test_two_solvers_each_get_their_binary():
    with sharer(3, 2) as s:
        assert s.trySetSolverValues(0, [T(0)])
        assert s.trySetSolverValues(1, [F(0)])
        s.trySendAssignment(0)
        s.trySendAssignment(1)
        a = s.addClause(-1, [F(0), mkLit(1)])
        b = s.addClause(-1, [mkLit(0), F(1)])
        s.execute()
        assert s.popAll(0) == [([F(0), mkLit(1)], a)]
        assert s.popAll(1) == [([mkLit(0), F(1)], b)]


def test_unset_values_are_unset_on_the_device():
    with sharer(2, 1) as s:
        assert s.trySetSolverValues(0, [T(1)])
        s.trySendAssignment(0)
        s.execute()
        assert s.popAll(0) == []
        s.unsetSolverValues(0, [T(1)])
        assert s.trySetSolverValues(0, [T(0)])
        s.trySendAssignment(0)
        cid = s.addClause(-1, [F(0), mkLit(1)])
        s.execute()
        # reported because x1 is no longer set
        assert s.popAll(0) == [([F(0), mkLit(1)], cid)]


def test_unary_clause_reported_as_unit():
    with sharer(2, 1) as s:
        assert s.trySetSolverValues(0, [T(1)])
        s.trySendAssignment(0)
        cid = s.addClause(-1, [mkLit(0)])
        s.execute()
        assert s.popAll(0) == [([mkLit(0)], cid)]
        assert s.getOneSolverStat(0, "reportedClausesUnit") == 1
        assert s.getOneSolverStat(0, "reportedClauses") == 1


def test_one_clause_reported_per_run():
    # the solver imports each unit clause and sends its new trail: only the
    # newest clause is unit under the next assignment
    n = 20
    with sharer(n, 1) as s:
        for i in range(n):
            cid = s.addClause(-1, [mkLit(i)])
            s.trySendAssignment(0)
            s.execute()
            assert s.popAll(0) == [([mkLit(i)], cid)]
            assert s.trySetSolverValues(0, [T(i)])
        assert s.getOneSolverStat(0, "reportedClauses") == n


def test_clauses_useful_for_earlier_assignments_are_reported():
    with sharer(4, 1) as s:
        a = s.addClause(-1, [F(0), mkLit(2)])
        b = s.addClause(-1, [F(1), mkLit(3)])
        assert s.trySetSolverValues(0, [T(0)])
        s.trySendAssignment(0)
        s.unsetSolverValues(0, [T(0)])
        assert s.trySetSolverValues(0, [T(1)])
        s.trySendAssignment(0)
        s.unsetSolverValues(0, [T(1)])
        assert s.getOneSolverStat(0, "reportedClauses") == 0
        s.execute()
        assert sorted(c for _, c in s.popAll(0)) == sorted([a, b])
        assert s.getOneSolverStat(0, "reportedClauses") == 2


def test_clause_counts_after_add_and_own_clause():
    with sharer(3, 1) as s:
        s.addClause(-1, [F(0), F(1)])
        s.gpuRun()
        assert s.getGlobalStat("gpuClauses") == 1
        assert s.getGlobalStat("gpuClauseLengthSum") == 2
    # a clause the solver learned itself is not handed back to it
    with sharer(3, 1) as s:
        assert s.trySetSolverValues(0, [T(0)])
        s.trySendAssignment(0)
        s.addClause(0, [F(0), F(1)])
        s.execute()
        assert s.getGlobalStat("gpuClauses") == 1
        assert s.getGlobalStat("gpuClauseLengthSum") == 2
        assert s.popAll(0) == []
        assert s.getOneSolverStat(0, "reportedClauses") == 0


def test_many_clauses_reported_all_at_once():
    n = 4000
    with sharer(n, 1) as s:
        assert s.trySetSolverValues(0, [T(i) for i in range(0, n, 3)])
        for i in range(n):
            s.addClause(-1, [mkLit(i)])
        s.trySendAssignment(0)
        s.execute()
        got = s.popAll(0)
        assert len(got) == n - (n + 2) // 3
        assert all(l[0] >> 1 != 0 and (l[0] >> 1) % 3 != 0 for l, _ in got)
        assert s.getOneSolverStat(0, "reportedClausesUnit") == len(got)


def test_solver_passes_many_assignments():
    with sharer(64, 3) as s:
        for i in range(32):
            s.addClause(-1, [F(2 * i), mkLit(2 * i + 1)])
            if i:
                s.unsetSolverValues(0, [T(2 * i - 2)])
            assert s.trySetSolverValues(0, [T(2 * i)])
            assert s.trySendAssignment(0) == i
        s.execute()
        got = s.popAll(0)
        assert len(got) == 32
        assert s.getOneSolverStat(0, "failuresToFindAssig") == 0
        assert s.getOneSolverStat(0, "reportedClausesBinary") == 32


def test_slot_window_exhaustion_and_deferred_unsets():
    with sharer(4, 1) as s:
        slots = s.slots_per_solver
        s.addClause(-1, [mkLit(3)])
        for i in range(slots):
            assert s.trySetSolverValues(0, [T(i % 3)])
            assert s.trySendAssignment(0) == i
        assert s.trySendAssignment(0) == -1
        assert not s.trySetSolverValues(0, [T(3)])
        assert s.getOneSolverStat(0, "failuresToFindAssig") == 2
        # an unset while full is deferred, and visible in the current assignment
        s.unsetSolverValues(0, [T(0)])
        cur = s.getCurrentAssignment(0, 4)
        assert list(cur) == [2, 0, 0, 2]
        s.execute()
        assert s.trySendAssignment(0) == slots
        assert len(s.popAll(0)) == 1  # {x3} unit in every assignment


def test_more_solvers_after_runs_keep_each_solvers_table_word():
    """Growing the solver count re-lays the assignment table out (1 -> 4
    words); solver 0's word keeps its values, so its next assignment, sent
    as the changes since the last one, is checked against all of them."""
    with sharer(4, 1) as s:
        a = s.addClause(-1, [F(0), mkLit(1)])
        assert s.trySetSolverValues(0, [T(0)])
        s.trySendAssignment(0)
        s.execute()
        assert s.popAll(0) == [([F(0), mkLit(1)], a)]
        s.setCpuSolverCount(3)
        assert list(s.getCurrentAssignment(0, 4)) == [0, 2, 2, 2]
        b = s.addClause(-1, [F(0), F(2), mkLit(3)])
        assert s.trySetSolverValues(0, [T(2)])  # x0 still true from before
        s.trySendAssignment(0)
        assert s.trySetSolverValues(2, [T(0), T(2)])
        s.trySendAssignment(2)
        s.execute()
        assert sorted(c for _, c in s.popAll(0)) == [a, b]
        assert s.popAll(1) == []
        assert sorted(c for _, c in s.popAll(2)) == [a, b]


def test_more_variables_after_runs():
    """The table grows in place when the variable count grows, and when a
    clause names a variable past it; earlier values stay."""
    with sharer(3, 1) as s:
        a = s.addClause(-1, [F(0), mkLit(1)])
        assert s.trySetSolverValues(0, [T(0)])
        s.trySendAssignment(0)
        s.execute()
        assert [c for _, c in s.popAll(0)] == [a]
        s.setVarCount(10)
        b = s.addClause(-1, [F(0), F(8), mkLit(9)])
        c = s.addClause(-1, [F(0), mkLit(20)])  # past the variable count
        assert s.trySetSolverValues(0, [T(8)])
        s.trySendAssignment(0)
        s.execute()
        assert sorted(x for _, x in s.popAll(0)) == [a, b, c]


def test_current_assignment_values():
    with sharer(5, 2) as s:
        s.trySetSolverValues(1, [T(0), F(2), T(4)])
        assert list(s.getCurrentAssignment(1, 5)) == [0, 2, 1, 2, 0]
        s.unsetSolverValues(1, [F(2)])
        assert list(s.getCurrentAssignment(1, 5)) == [0, 2, 2, 2, 0]
        assert list(s.getCurrentAssignment(0, 5)) == [2] * 5
        # a variable past the sharer's count: the buffer still gets exactly
        # one byte per counted variable
        assert s.trySetSolverValues(0, [T(1), F(9)])
        assert list(s.getCurrentAssignment(0, 5)) == [2, 0, 2, 2, 2]
        s.unsetSolverValues(0, [F(9)])
        assert list(s.getCurrentAssignment(0, 5)) == [2, 0, 2, 2, 2]


def test_own_clause_not_reported_back():
    with sharer(3, 2) as s:
        for solver in (0, 1):
            assert s.trySetSolverValues(solver, [T(0)])
            s.trySendAssignment(solver)
        cid = s.addClause(0, [F(0), mkLit(1)])
        s.execute()
        assert s.popAll(0) == []
        assert s.popAll(1) == [([F(0), mkLit(1)], cid)]


def test_reduce_db_keeps_active_clauses():
    with sharer(5, 1) as s:
        s.addClause(-1, [F(1), mkLit(3), mkLit(4)])
        used = s.addClause(-1, [F(0), mkLit(2), mkLit(4)])
        assert s.trySetSolverValues(0, [T(0), F(4)])
        s.trySendAssignment(0)
        s.execute()
        assert s.getGlobalStat("gpuClauses") == 2
        assert s.getGlobalStat("gpuClauseLengthSum") == 6
        assert [c for _, c in s.popAll(0)] == [used]
        s.reduceDb()
        assert s.getGlobalStat("gpuClauses") == 1
        assert s.getGlobalStat("gpuClauseLengthSum") == 3
        assert s.getGlobalStat("gpuReduceDbs") == 1
        assert s.getAddedClauseCountAtLastReduceDb() == 2
        assert s.trySetSolverValues(0, [T(1)])
        s.trySendAssignment(0)
        s.execute()
        assert all(c == used for _, c in s.popAll(0))
        rep = s.getOneSolverStat(0, "reportedClauses")
        s.addClause(-1, [F(0), F(3)])
        s.trySendAssignment(0)
        s.execute()
        s.popAll(0)
        assert s.getOneSolverStat(0, "reportedClauses") >= rep + 1
        assert s.getAddedClauseCount() == 3


def test_reduce_db_with_a_run_in_flight():
    """reduceDb delivers the run in flight before rows move
    (GpuClauseSharerImpl.cu:105-110): its reports are the clauses they
    named, though the reduction then removes some of them."""
    with sharer(6, 1) as s:
        ids = [s.addClause(-1, [F(0), mkLit(v), mkLit(v + 1)]) for v in (1, 2, 3, 4)]
        assert s.trySetSolverValues(0, [T(0), F(2), F(3)])
        s.trySendAssignment(0)
        s.gpuRun()  # the sweep is in flight
        s.reduceDb()
        got = {c: l for l, c in s.popAll(0)}
        # x0 true, x2 and x3 false: (~x0 v x1 v x2) and (~x0 v x3 v x4) are
        # unit, (~x0 v x2 v x3) is falsified, (~x0 v x4 v x5) has two
        # undefined literals
        assert sorted(got) == ids[:3]
        for i, v in enumerate((1, 2, 3)):
            assert got[ids[i]] == [F(0), mkLit(v), mkLit(v + 1)]
        assert s.getGlobalStat("gpuReduceDbs") == 1


def test_activity_init_bump_decay():
    # ClauseActivityLbdTest.cu:59-95 with decay 0.5: a new clause starts at the
    # increment after one decay (2.0); a report bumps it by the increment.
    with sharer(2, 1, clause_activity_decay=0.5) as s:
        cid = s.addClause(-1, [mkLit(0)])
        s.gpuRun()
        act, lbd = s.debug_clause(cid)
        assert abs(act - 2.0) < 1e-6 and lbd == 1
        s.trySendAssignment(0)  # x0 undefined: {x0} is unit
        s.execute()
        act, _ = s.debug_clause(cid)
        assert abs(act - 4.0) < 1e-6
        s.addClause(-1, [mkLit(1)])  # one more decay: increment 4.0
        s.trySendAssignment(0)
        s.execute()
        act, _ = s.debug_clause(cid)
        assert abs(act - 8.0) < 1e-6


def test_stat_names_and_counts():
    with sharer(3, 1) as s:
        assert s.getGlobalStatCount() == 13
        assert s.getOneSolverStatCount() == 6
        assert s.getGlobalStatName(0) == "gpuClauses"
        assert s.getGlobalStatName(12) == "timeSpentReduceGpuDb"
        assert s.getOneSolverStatName(5) == "reportedClausesBinary"
        free, total = s.getGpuMemInfo()
        assert 0 < free <= total


def test_write_clauses_in_cnf(tmp_path):
    with sharer(4, 1) as s:
        s.addClause(-1, [F(0), mkLit(3)])
        s.addClause(-1, [mkLit(1)])
        s.gpuRun()
        p = tmp_path / "db.cnf"
        s.writeClausesInCnf(str(p))
        assert p.read_text().splitlines() == ["p cnf 4 2", "2 0", "-1 4 0"]


def test_out_of_device_memory_halves_the_db(tmp_path, monkeypatch):
    """An append that does not fit makes gpuRun reduce the DB and carry on,
    and hasRunOutOfGpuMemoryOnce() turns true (GpuRunner.cu:243-260).  The
    condition is forced from the host by capping the clause DB's device
    bytes (GPUSHARESAT_DB_LIMIT_BYTES, read when the engine is created).
    The append fails while a sweep with hits is in flight: those hits are
    delivered as the clauses they named before the reduction moved rows
    (GpuRunner.cu:251-253).  Afterwards the device holds exactly the host's
    clauses, those of the failed batch in the always-kept lengths 1-2
    included: an assignment's reports equal the clauses of the written CNF
    that it makes falsified or unit."""
    from collections import Counter
    monkeypatch.setenv("GPUSHARESAT_DB_LIMIT_BYTES", str(256 << 10))
    rng = np.random.default_rng(5)
    V = 400

    def add(s, n, lo, hi):
        out = []
        for _ in range(n):
            vs = rng.choice(V, size=int(rng.integers(lo, hi)), replace=False)
            out.append([mkLit(int(v), bool(rng.integers(2))) for v in vs])
            s.addClause(-1, out[-1])
        return out

    def assignment():
        return {int(v): bool(rng.integers(2)) for v in rng.choice(V, size=V // 2, replace=False)}

    def useful(clauses, assig):
        return Counter(tuple(sorted(c)) for c in clauses if _useful(c, assig))

    with sharer(V, 1) as s:
        first = add(s, 2000, 3, 7) + add(s, 300, 1, 3)
        s.execute()
        assert not s.hasRunOutOfGpuMemoryOnce()
        a1 = assignment()
        trail = [mkLit(v, not val) for v, val in a1.items()]
        assert s.trySetSolverValues(0, trail)
        assert s.trySendAssignment(0) >= 0
        s.gpuRun()  # a sweep of the first clauses against a1 is in flight
        add(s, 12000, 3, 7)  # ~216 KB more of clauses: past the cap
        add(s, 2000, 1, 3)
        s.gpuRun()  # a1's hits collected, the append fails, they are delivered, the DB reduced
        assert s.hasRunOutOfGpuMemoryOnce()
        assert s.getGlobalStat("gpuReduceDbs") == 1
        got1 = Counter(tuple(sorted(l)) for l, _ in s.popAll(0))
        exp1 = useful(first, a1)
        assert sum(exp1.values()) > 0 and got1 == exp1
        kept = s.getGlobalStat("gpuClauses")
        assert 2300 <= kept < 16300
        s.unsetSolverValues(0, trail)
        a2 = assignment()
        assert s.trySetSolverValues(0, [mkLit(v, not val) for v, val in a2.items()])
        assert s.trySendAssignment(0) >= 0
        s.execute()
        got = Counter(tuple(sorted(l)) for l, _ in s.popAll(0))
        p = tmp_path / "db.cnf"
        s.writeClausesInCnf(str(p))
        lines = p.read_text().splitlines()
        assert lines[0].split()[3] == str(kept)
        db = [[mkLit(abs(int(x)) - 1, int(x) < 0) for x in line.split()[:-1]] for line in lines[1:]]
        exp = useful(db, a2)
        assert sum(exp.values()) > 0
        assert got == exp


def _db_matches_cnf(s, tmp_path, V, rng, solver=0):
    """The device holds exactly the clauses of the written CNF: one
    assignment's reports equal the CNF clauses it makes falsified or unit
    (for a solver with no reports from earlier assignments still held back,
    Reported.cu:97-158)."""
    from collections import Counter
    p = tmp_path / "db.cnf"
    s.writeClausesInCnf(str(p))
    lines = p.read_text().splitlines()
    db = [[mkLit(abs(int(x)) - 1, int(x) < 0) for x in line.split()[:-1]] for line in lines[1:]]
    assig = {int(v): bool(rng.integers(2)) for v in rng.choice(V, size=V // 2, replace=False)}
    s.unsetSolverValues(solver, [mkLit(v) for v in range(V)])
    assert s.trySetSolverValues(solver, [mkLit(v, not val) for v, val in assig.items()])
    assert s.trySendAssignment(solver) >= 0
    s.execute()
    got = Counter(tuple(sorted(l)) for l, _ in s.popAll(solver))
    exp = Counter(tuple(sorted(c)) for c in db if _useful(c, assig))
    assert got == exp
    return len(db)


def test_out_of_memory_reduces_until_the_db_fits(tmp_path, monkeypatch):
    """An append far past the DB byte cap: one halving still does not fit, so
    the reduction repeats (as the reference's reduceDb loop) until the upload
    fits; gpuRun carries on and the device holds exactly the mirror."""
    monkeypatch.setenv("GPUSHARESAT_DB_LIMIT_BYTES", str(128 << 10))
    rng = np.random.default_rng(8)
    V = 500
    with sharer(V, 1) as s:
        for _ in range(16000):  # ~320 KB of literals of lengths 3-7
            vs = rng.choice(V, size=int(rng.integers(3, 8)), replace=False)
            s.addClause(-1, [mkLit(int(v), bool(rng.integers(2))) for v in vs])
        s.gpuRun()
        assert s.hasRunOutOfGpuMemoryOnce()
        kept = s.getGlobalStat("gpuClauses")
        assert 0 < kept <= 16000 // 4 + 200  # at least two halvings
        assert _db_matches_cnf(s, tmp_path, V, rng) == kept


def test_out_of_memory_with_nothing_left_to_reduce(tmp_path, monkeypatch):
    """Binary clauses are never reduced (Clauses.cu:255): when they alone do
    not fit, the buckets the device cannot hold are emptied on both sides and
    the sharer keeps working with device and mirror equal."""
    monkeypatch.setenv("GPUSHARESAT_DB_LIMIT_BYTES", str(16 << 10))
    rng = np.random.default_rng(9)
    V = 300
    def add(n, size):
        for _ in range(n):
            s.addClause(-1, [mkLit(int(v), bool(rng.integers(2)))
                             for v in rng.choice(V, size=size, replace=False)])

    with sharer(V, 1) as s:
        add(100, 1)
        s.gpuRun()
        assert not s.hasRunOutOfGpuMemoryOnce()
        add(4000, 2)  # 32 KB of binary clauses
        s.gpuRun()
        assert s.hasRunOutOfGpuMemoryOnce()
        # the binary bucket could not be held; the unary one is untouched
        assert s.getGlobalStat("gpuClauses") == 100
        assert _db_matches_cnf(s, tmp_path, V, rng) == 100
        add(50, 2)  # fits again
        s.gpuRun()
        assert s.getGlobalStat("gpuClauses") == 150
        assert _db_matches_cnf(s, tmp_path, V, rng) == 150


def test_pageable_host_memory_option():
    """maxPageLockedMemory = 1 byte (GpuClauseSharer.h:47): every host copy
    endpoint is pageable (CorrespArr.cu:212-235); reports are unchanged."""
    with sharer(4, 2, max_page_locked_memory=1) as s:
        a = s.addClause(-1, [F(0), mkLit(1)])
        b = s.addClause(-1, [mkLit(0), F(1), mkLit(2)])
        assert s.trySetSolverValues(0, [T(0)])
        assert s.trySetSolverValues(1, [F(0), T(1)])
        s.trySendAssignment(0)
        s.trySendAssignment(1)
        s.execute()
        assert s.popAll(0) == [([F(0), mkLit(1)], a)]
        assert s.popAll(1) == [([mkLit(0), F(1), mkLit(2)], b)]


def test_rejected_clauses_and_bad_ids():
    from gpusharesat_amd._lib import GssError
    with sharer(4, 1) as s:
        assert s.addClause(-1, []) == -1
        assert s.addClause(-1, [0] * (s.max_clause_size + 1)) == -1
        with pytest.raises(GssError):
            s.trySendAssignment(5)
        with pytest.raises(GssError):
            s.addClause(3, [0])
        # literals past the device's variable range (2^24 - 1 variables) are
        # refused where they enter, before any state changes
        assert s.addClause(-1, [mkLit((1 << 24) - 1)]) == -1
        with pytest.raises(GssError):
            s.trySetSolverValues(0, [mkLit((1 << 24) - 1)])
        with pytest.raises(GssError):
            s.unsetSolverValues(0, [mkLit((1 << 24) - 1)])
        cid = s.addClause(-1, [mkLit((1 << 24) - 2)])  # the largest variable is fine
        assert cid >= 0
        s.execute()
        assert s.getGlobalStat("gpuClauses") == 1


def _clause_hits(clauses, snapshots):
    """Oracle: for each snapshot (var values 0/1/2), the set of clause indices
    that are falsified or unit."""
    V = len(snapshots[0])
    A = len(snapshots)
    W = (A + 63) // 64
    d = np.zeros((V, W), np.uint64)
    t = np.zeros((V, W), np.uint64)
    for a, vals in enumerate(snapshots):
        bit = np.uint64(1 << (a % 64))
        d[vals != 2, a // 64] |= bit
        t[vals == 0, a // 64] |= bit
    off = np.zeros(len(clauses) + 1, np.int64)
    off[1:] = np.cumsum([len(c) for c in clauses])
    lits = np.concatenate([np.asarray(c, np.uint32) for c in clauses])
    valid = np.zeros(W, np.uint64)
    for a in range(A):
        valid[a // 64] |= np.uint64(1 << (a % 64))
    out = [set() for _ in range(A)]
    for h in orc.sweep(off, lits, d, t, valid):
        m = int(h["false_mask"]) | int(h["unit_mask"])
        for b in range(64):
            if m >> b & 1:
                out[int(h["word"]) * 64 + b].add(int(h["clause"]))
    return out


@pytest.mark.parametrize("solvers,seed,screen,rounds", [
    (1, 0, "0", 12), (3, 1, "0", 12), (5, 2, "0", 12), (16, 3, "0", 12), (3, 4, "1", 12),
    (16, 5, "1", 12), (3, 6, "auto:8", 40)])
def test_snapshot_semantics_match_oracle(monkeypatch, solvers, seed, screen, rounds):
    """Each solver mutates its assignment (set / unset), sends 0-3 snapshots
    per run, and pops everything after two runs; the popped clause ids must be
    exactly the union of the oracle's hits over its snapshots of that run.
    With the runs screened against the table's aggregate (GPUSHARESAT_SCREEN
    1), not (0), and chosen by their times (auto, the default, here with
    the other form for 4 of every 8 runs instead of 256, so that the forms
    alternate many times)."""
    monkeypatch.setenv("GPUSHARESAT_SCREEN", screen)
    rng = np.random.default_rng(seed)
    V = 60
    with sharer(V, solvers) as s:
        clauses = []
        ids = []
        state = [np.full(V, 2, np.int64) for _ in range(solvers)]
        for rnd in range(rounds):
            for _ in range(int(rng.integers(5, 40))):
                L = int(rng.integers(1, 6))
                c = [int(x) for x in rng.integers(0, 2 * V, size=L)]
                clauses.append(c)
                ids.append(s.addClause(-1, c))
            snaps = [[] for _ in range(solvers)]
            for so in range(solvers):
                for _ in range(int(rng.integers(0, 4))):
                    setv = rng.choice(V, size=int(rng.integers(0, 12)), replace=False)
                    lits = [mkLit(int(v), bool(rng.integers(0, 2))) for v in setv]
                    unset = [int(v) for v in rng.choice(V, size=int(rng.integers(0, 8)),
                                                         replace=False)]
                    s.unsetSolverValues(so, [mkLit(v) for v in unset])
                    for v in unset:
                        state[so][v] = 2
                    assert s.trySetSolverValues(so, lits)
                    for l in lits:
                        state[so][l >> 1] = 1 if l & 1 else 0
                    assert list(s.getCurrentAssignment(so, V)) == list(state[so])
                    assert s.trySendAssignment(so) >= 0
                    snaps[so].append(state[so].copy())
            s.execute()
            for so in range(solvers):
                got = sorted(cid for _, cid in s.popAll(so))
                exp = set()
                if snaps[so]:
                    for hs in _clause_hits(clauses, snaps[so]):
                        exp |= hs
                assert got == sorted(ids[i] for i in exp), (rnd, so)


def _useful(lits, assig):
    """Falsified or unit under a var -> bool assignment (unset = undefined)."""
    undef = 0
    for x in lits:
        v, neg = x >> 1, x & 1
        if v not in assig:
            undef += 1
        elif assig[v] != bool(neg):
            return False  # a true literal
    return undef <= 1


def test_solver_threads_and_gpu_thread_concurrently():
    """The threading contract (GpuClauseSharer.h:78-158): each solver thread
    sets, unsets, sends, pops and adds clauses while one thread loops
    gpuRun.  Every clause handed to a solver is falsified or unit under one
    of the assignments that solver sent, and nothing is lost or crashes."""
    import threading
    V, S, ROUNDS = 40, 4, 150
    rng0 = np.random.default_rng(17)
    with sharer(V, S) as s:
        for _ in range(300):
            k = int(rng0.integers(2, 5))
            vs = rng0.choice(V, size=k, replace=False)
            s.addClause(-1, [mkLit(int(v), bool(rng0.integers(2))) for v in vs])
        stop = threading.Event()
        errors = []
        sent = [[] for _ in range(S)]
        got = [[] for _ in range(S)]

        def gpu():
            try:
                while not stop.is_set():
                    s.gpuRun()
            except Exception as ex:  # reported after the join
                errors.append(ex)

        def solver(sid):
            try:
                rng = np.random.default_rng(100 + sid)
                cur = {}
                for _ in range(ROUNDS):
                    if cur:
                        s.unsetSolverValues(sid, [mkLit(v, not val) for v, val in cur.items()])
                        cur = {}
                    vs = rng.choice(V, size=int(0.6 * V), replace=False)
                    new = {int(v): bool(rng.integers(2)) for v in vs}
                    if s.trySetSolverValues(sid, [mkLit(v, not val) for v, val in new.items()]):
                        cur = new
                    if s.trySendAssignment(sid) >= 0:
                        sent[sid].append(dict(cur))
                    got[sid].extend(l for l, _ in s.popAll(sid))
                    if rng.random() < 0.1:
                        vs = rng.choice(V, size=3, replace=False)
                        s.addClause(sid, [mkLit(int(v), bool(rng.integers(2))) for v in vs])
            except Exception as ex:
                errors.append(ex)

        g = threading.Thread(target=gpu, daemon=True)
        g.start()
        ts = [threading.Thread(target=solver, args=(i,), daemon=True) for i in range(S)]
        for t in ts:
            t.start()
        for t in ts:
            t.join(timeout=60)
        stop.set()
        g.join(timeout=60)
        assert not g.is_alive() and not any(t.is_alive() for t in ts)
        assert not errors, errors
        s.execute()
        for sid in range(S):
            got[sid].extend(l for l, _ in s.popAll(sid))
        print("sent", [len(x) for x in sent], "reported", [len(x) for x in got], flush=True)
        assert sum(len(x) for x in sent) > 0 and sum(len(x) for x in got) > 0
        for sid in range(S):
            for lits in got[sid]:
                assert any(_useful(lits, a) for a in sent[sid]), (sid, lits)


def test_device_activity_policy_keeps_active_clauses(monkeypatch):
    """GPUSHARESAT_REDUCE_POLICY=device: reduceDb runs the reference's rule
    (keep activity >= the approximate median, Clauses.cu:391-525) on the
    device over a tagged DB (gss_res_reduce); the reported clause stays."""
    monkeypatch.setenv("GPUSHARESAT_REDUCE_POLICY", "device")
    with sharer(5, 1) as s:
        s.addClause(-1, [F(1), mkLit(3), mkLit(4)])
        used = s.addClause(-1, [F(0), mkLit(2), mkLit(4)])
        assert s.trySetSolverValues(0, [T(0), F(4)])
        s.trySendAssignment(0)
        s.execute()
        assert [c for _, c in s.popAll(0)] == [used]
        s.reduceDb()
        assert s.getGlobalStat("gpuClauses") == 1
        assert s.getGlobalStat("gpuClauseLengthSum") == 3
        s.unsetSolverValues(0, [T(0), F(4)])
        assert s.trySetSolverValues(0, [T(0), F(2), F(4)])
        s.trySendAssignment(0)
        s.execute()
        assert [c for _, c in s.popAll(0)] == [used]


@pytest.mark.parametrize("policy", ["device", "learned"])
def test_device_policies_reduce_and_agree_with_the_mirror(tmp_path, monkeypatch, policy):
    """The device policies through reduceDb: about half of the clauses of
    length >= 3 go, every clause of length <= 2 stays (Clauses.cu:255), and
    the device holds exactly the mirror's clauses (reports of a fresh
    assignment = the clauses of the written CNF it makes falsified or unit).
    The learned policy trains on the run's own records every 16 runs."""
    monkeypatch.setenv("GPUSHARESAT_REDUCE_POLICY", policy)
    rng = np.random.default_rng(31)
    V = 300
    with sharer(V, 3) as s:  # solvers 0 and 1 send; solver 2 checks the DB at the end
        shorts = 0
        for _ in range(3000):
            n = int(rng.integers(1, 9))
            shorts += n <= 2
            s.addClause(-1, [mkLit(int(v), bool(rng.integers(2)))
                             for v in rng.choice(V, size=n, replace=False)])
        for r in range(40):  # runs with assignments: records feed the policy state
            for so in range(2):
                s.unsetSolverValues(so, [mkLit(v) for v in range(V)])
                assig = rng.choice(V, size=V // 2, replace=False)
                assert s.trySetSolverValues(so, [mkLit(int(v), bool(rng.integers(2)))
                                                 for v in assig])
                s.trySendAssignment(so)
            s.gpuRun()
            s.popAll(0)
            s.popAll(1)
        before = s.getGlobalStat("gpuClauses")
        s.reduceDb()
        after = s.getGlobalStat("gpuClauses")
        longs = before - shorts
        assert shorts + 0.3 * longs <= after <= shorts + 0.7 * longs, (before, after, shorts)
        p = tmp_path / "db.cnf"
        s.writeClausesInCnf(str(p))
        lens = [len(line.split()) - 1 for line in p.read_text().splitlines()[1:]]
        assert sum(1 for n in lens if n <= 2) == shorts
        assert _db_matches_cnf(s, tmp_path, V, rng, solver=2) == after


def test_device_policy_out_of_memory_reduces_and_retries(tmp_path, monkeypatch):
    """Device policy with a DB byte cap: the append that does not fit makes
    the device reduce and the batch is added again, until it fits; device
    and mirror agree afterwards."""
    monkeypatch.setenv("GPUSHARESAT_REDUCE_POLICY", "device")
    monkeypatch.setenv("GPUSHARESAT_DB_LIMIT_BYTES", str(160 << 10))
    rng = np.random.default_rng(33)
    V = 400
    with sharer(V, 1) as s:
        for _ in range(4000):
            vs = rng.choice(V, size=int(rng.integers(3, 8)), replace=False)
            s.addClause(-1, [mkLit(int(v), bool(rng.integers(2))) for v in vs])
        s.gpuRun()
        assert not s.hasRunOutOfGpuMemoryOnce()
        for _ in range(4000):
            vs = rng.choice(V, size=int(rng.integers(3, 8)), replace=False)
            s.addClause(-1, [mkLit(int(v), bool(rng.integers(2))) for v in vs])
        s.gpuRun()
        assert s.hasRunOutOfGpuMemoryOnce()
        kept = s.getGlobalStat("gpuClauses")
        assert 4000 < kept < 8000  # the new batch and what was kept of the first
        assert _db_matches_cnf(s, tmp_path, V, rng) == kept
