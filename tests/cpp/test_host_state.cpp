// CPU tests of the sharer's host state (gpusharesat_amd/csrc/host_state.cpp):
// the reference's known answers for assignment windows, clause activity and
// per-solver delivery, restated.  Built with g++ and run by
// tests/test_host_state.py; prints one line per failed check and exits
// non-zero if any failed.
//
//   GpuSolverTest.cu:78-181   assignment encodings, carry-forward of completed
//                             slots, window exhaustion
//   ClauseActivityLbdTest.cu  init/bump/decay, approxNthAct intervals,
//                             act-only removal threshold, rescale
//   GpuSolverTest.cu:1037-1050, Reported.cu:105-158
//                             batch pop order, dedup, release, own clauses
//   ConcurrentQueueTest.cc:29-70
//                             batches handed out in publication order
#include <cmath>
#include <cstdio>
#include <cstring>
#include <set>
#include <vector>

#include "sharer_impl.h"

using namespace GpuShare;

static int failures = 0, checks = 0;
#define CHECK(cond)                                                        \
    do {                                                                   \
        checks++;                                                          \
        if (!(cond)) {                                                     \
            failures++;                                                    \
            std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond);    \
        }                                                                  \
    } while (0)

static int lit(int v, bool neg = false) { return 2 * v + (neg ? 1 : 0); }

static const gss_var_update *find(const std::vector<gss_var_update> &u, int var) {
    for (const auto &x : u)
        if ((int)x.var == var) return &x;
    return nullptr;
}

// testAssigsTwoAssignments (GpuSolverTest.cu:125-155)
static void two_assignments() {
    SolverWindow w(64);
    w.set_var_count(2);
    w.set_var(0, kValTrue);
    w.set_var(1, kValFalse);
    CHECK(w.done() == 0);
    w.set_var(0, kValUndef);
    CHECK(w.done() == 1);
    std::vector<gss_var_update> ups;
    SolverWindow::Taken t = w.take(ups, 3);
    CHECK(t.start_id == 0 && t.count == 2 && t.valid == 0b11);
    const gss_var_update *v0 = find(ups, 0), *v1 = find(ups, 1);
    CHECK(v0 && v1 && v0->word == 3);
    CHECK((v0->is_def & 1) == 1 && (v0->is_true & 1) == 1);  // slot 0: x0 true
    CHECK((v0->is_def & 2) == 0);                            // slot 1: x0 undefined
    CHECK((v1->is_def & 3) == 3 && (v1->is_true & 3) == 0);  // x1 false in both
    // roll-forward to the slot after the last completed one (current id 2)
    CHECK(t.last_pos == 2);
}

// testAssigsTwoSolvers (GpuSolverTest.cu:78-123): independent windows
static void two_solvers() {
    SolverWindow a(64), b(64);
    a.set_var_count(2);
    b.set_var_count(2);
    a.set_var(0, kValFalse);
    a.set_var(1, kValUndef);
    a.done();
    b.set_var(0, kValUndef);
    b.set_var(1, kValTrue);
    b.done();
    std::vector<gss_var_update> ua, ub;
    a.take(ua, 0);
    b.take(ub, 1);
    CHECK((find(ua, 0)->is_def & 1) == 1 && (find(ua, 0)->is_true & 1) == 0);
    CHECK((find(ua, 1)->is_def & 1) == 0);
    CHECK((find(ub, 0)->is_def & 1) == 0);
    CHECK((find(ub, 1)->is_def & 1) == 1 && (find(ub, 1)->is_true & 1) == 1);
}

// testManyAssignments (GpuSolverTest.cu:157-209): exhaustion, then a new
// window whose first slot is the next id modulo the slot count
static void many_assignments(int slots) {
    SolverWindow w(slots);
    w.set_var_count(slots);
    for (int i = 0; i < slots; i++) {
        CHECK(w.available());
        w.set_var(i, i % 2 == 0 ? kValTrue : kValFalse);
        CHECK(w.done() == i);
    }
    CHECK(!w.available());
    std::vector<gss_var_update> ups;
    SolverWindow::Taken t = w.take(ups, 0);
    CHECK(t.count == slots);
    CHECK(t.valid == (slots == 64 ? ~0ull : ((1ull << slots) - 1)));
    CHECK(t.last_pos == slots - 1);  // full window: the last completed slot
    // var i was set for assignment i: slots before i see it undefined
    const gss_var_update *u5 = find(ups, 5);
    CHECK(u5 && (u5->is_def & ((1ull << 5) - 1)) == 0 && (u5->is_def >> 5 & 1) == 1);
    CHECK(w.available());
    w.set_var(7, kValUndef);
    CHECK(w.done() == slots);
    ups.clear();
    t = w.take(ups, 0);
    CHECK(t.start_id == slots && t.count == 1 && t.valid == 1);  // slot = id % slots = 0
    const gss_var_update *u7 = find(ups, 7);
    CHECK(u7 && (u7->is_def & 1) == 0);
    uint8_t cur[64];
    w.current_assignment(cur, 8);
    CHECK(cur[7] == kValUndef && cur[6] == kValTrue && cur[5] == kValFalse);
    // exactly n values: variables the window never saw are undefined, and
    // nothing is written past n
    std::vector<uint8_t> big((size_t)slots + 4, 0xAB);
    w.current_assignment(big.data(), slots + 2);
    CHECK(big[6] == kValTrue && big[slots] == kValUndef && big[slots + 1] == kValUndef &&
          big[slots + 2] == 0xAB);
    uint8_t small[4];
    std::memset(small, 0xAB, sizeof small);
    w.current_assignment(small, 3);
    CHECK(small[2] == kValTrue && small[1] == kValFalse && small[3] == 0xAB);
}

// carry-forward: a value changed after an assignment was completed keeps the
// old value in the completed slot (Assigs.cu:172-181)
static void carry_forward() {
    SolverWindow w(64);
    w.set_var_count(4);
    w.set_var(2, kValTrue);
    std::vector<gss_var_update> ups;
    w.take(ups, 0);  // x2 = true everywhere on the device from now on
    w.done();        // assignment 0 sees x2 true
    w.set_var(2, kValFalse);
    w.done();        // assignment 1 sees x2 false
    ups.clear();
    SolverWindow::Taken t = w.take(ups, 0);
    const gss_var_update *u = find(ups, 2);
    CHECK(t.valid == 0b11);
    CHECK(u && (u->is_def & 3) == 3 && (u->is_true & 3) == 1);
}

static void mirror_with(ClauseMirror &m, const std::vector<std::pair<int, int>> &lbd_act,
                        std::vector<long> *ids = nullptr) {
    int v = 0;
    for (auto la : lbd_act) {
        std::vector<int> l;
        for (int j = 0; j < la.first; j++) l.push_back(lit(v++));
        long id = m.stage(l.data(), la.first, la.first);
        if (ids) ids->push_back(id);
        ClauseMirror::Batch b;
        m.commit_staged(b);
        for (int k = 0; k < la.second - 1; k++) m.bump(la.first, b.rows[0]);
    }
}

static float act_of(const ClauseMirror &m, long id) {
    float a = -1;
    int lbd;
    m.find(id, a, lbd);
    return a;
}

// ClauseActivityLbdTest.cu:59-95
static void activity_basics() {
    std::vector<long> ids;
    ClauseMirror m1(1.0f);
    mirror_with(m1, {{1, 1}}, &ids);
    CHECK(std::fabs(act_of(m1, ids[0]) - 1.0f) < 0.01f);
    ids.clear();
    ClauseMirror m2(1.0f);
    mirror_with(m2, {{1, 2}}, &ids);
    CHECK(std::fabs(act_of(m2, ids[0]) - 2.0f) < 0.01f);
    // decay 0.5: a new clause starts at 2, one more decay, bump -> 6
    ClauseMirror m3(0.5f);
    int l0 = lit(0);
    long id = m3.stage(&l0, 1, 1);
    ClauseMirror::Batch b;
    m3.commit_staged(b);
    CHECK(std::fabs(act_of(m3, id) - 2.0f) < 0.01f);
    int l1 = lit(1);
    m3.stage(&l1, 1, 1);  // committing it decays once more: increment 4
    m3.commit_staged(b);
    m3.bump(1, 0);
    CHECK(std::fabs(act_of(m3, id) - 6.0f) < 0.01f);
}

// testApproxNthLargestAct* (ClauseActivityLbdTest.cu:163-192): activities given
// as bump counts with decay 1, all clauses of lbd 3, range [3, 4)
static void approx_nth() {
    auto check = [](std::vector<int> acts, float lo, float hi, long n) {
        std::vector<std::pair<int, int>> la;
        for (int a : acts) la.push_back({3, a});
        ClauseMirror m(1.0f);
        mirror_with(m, la);
        const float v = m.approx_nth_activity(3, 4, n);
        CHECK(v > lo && v <= hi);
        if (!(v > lo && v <= hi)) std::printf("  got %f, want (%f, %f]\n", v, lo, hi);
    };
    check({1, 2, 3}, 2.0f, 2.1f, 2);
    check({1, 2, 3}, -0.1f, 0.1f, 0);
    check({1, 2, 3}, 3.0f, 3.1f, 3);
    check({1, 1, 1, 3, 3, 100}, 1.0f, 1.1f, 3);
    check({1, 1, 1, 1, 10}, 1.0f, 1.1f, 2);
}

// testActOnly (ClauseActivityLbdTest.cu:224-229): act-only threshold over all
// lbds in [0, max) with n = count / 2
static void act_only_threshold() {
    ClauseMirror m(1.0f);
    mirror_with(m, {{2, 1}, {8, 5}});
    const float v = m.approx_nth_activity(0, 100, m.clause_count() / 2);
    CHECK(v > 1.0f && v <= 5.0f);
}

// testActOnlyHugeDifferences (ClauseActivityLbdTest.cu:231-253)
static void act_only_huge_differences() {
    ClauseMirror m(0.5f);
    std::vector<int> l = {lit(0), lit(1), lit(2)};
    m.stage(l.data(), 3, 3);  // act 2
    ClauseMirror::Batch b;
    m.commit_staged(b);
    m.stage(l.data(), 3, 3);  // act 4 (bumped once)
    m.commit_staged(b);
    m.bump(3, b.rows[0]);
    // Decay without adding clauses: activities of new clauses grow; the
    // reference calls decayClauseAct() directly, here each committed clause
    // decays once, so commit (1e19 / 100) / log(2) clauses of length 1.
    const int c = (int)(std::log(1e19 / 100) / std::log(1 / 0.5));
    int one = lit(9);
    for (int i = 0; i < c - 1; i++) m.stage(&one, 1, 1);
    m.commit_staged(b);
    m.stage(l.data(), 3, 3);
    m.commit_staged(b);
    // among the lbd-3 clauses the smallest is 2
    const float v = m.approx_nth_activity(3, 4, 1);
    CHECK(v > 2.0f && v <= 4.0f);
}

// testRescale (ClauseActivityLbdTest.cu:255-273)
static void rescale() {
    ClauseMirror m(1.0f / (1e19f * 10.0f));
    int l0 = lit(0);
    long id = m.stage(&l0, 1, 1);
    ClauseMirror::Batch b;
    m.commit_staged(b);
    CHECK(std::fabs(act_of(m, id) - 10.0f) < 0.1f);
    int l1 = lit(1);
    m.stage(&l1, 1, 1);  // one decay (increment 1e21 before the rescale)
    m.commit_staged(b);
    // the new clause started above the limit: everything was rescaled
    const float a_after = act_of(m, id);
    m.bump(1, 0);
    CHECK(std::fabs(act_of(m, id) - (a_after + m.act_incr())) < 0.1f);
}

// reduce: lengths 1-2 are never removed; the least active half goes
static void reduce_keeps_active() {
    std::vector<long> ids;
    ClauseMirror m(1.0f);
    mirror_with(m, {{3, 1}, {3, 5}, {2, 1}, {4, 1}, {4, 9}}, &ids);
    std::vector<int> changed = m.reduce(100);
    float a;
    int lbd;
    CHECK(!m.find(ids[0], a, lbd));  // act 1, length 3: removed
    CHECK(m.find(ids[1], a, lbd));   // act 5: kept
    CHECK(m.find(ids[2], a, lbd));   // length 2: kept
    CHECK(!m.find(ids[3], a, lbd));
    CHECK(m.find(ids[4], a, lbd));
    CHECK(m.clause_count() == 3 && m.length_sum() == 3 + 2 + 4);
    CHECK(changed.size() == 2);
    // rows are compacted in order
    CHECK(m.meta(3, 0).id == ids[1] && m.row_lits(3, 0)[0] == (uint32_t)lit(3));
}

static ClauseBatch *batch(ReportQueue &q, long start, long count,
                          std::vector<std::vector<int>> cls, long first_id) {
    ClauseBatch *b = q.get_new();
    b->start_assig = start;
    b->assig_count = count;
    long id = first_id;
    for (auto &c : cls) {
        b->clauses.push_back({id++, (int)b->lits.size()});
        b->lits.insert(b->lits.end(), c.begin(), c.end());
    }
    return b;
}

// Reported::popReportedClause semantics (Reported.cu:105-158)
static void report_queue() {
    std::atomic<unsigned long> st[6];
    for (auto &s : st) s = 0;
    ReportQueue q;
    int *lits;
    int count;
    long id;
    CHECK(!q.pop(lits, count, id, st));
    q.assig_was_sent(0);
    q.publish(batch(q, 0, 1, {{lit(1)}, {lit(2), lit(3)}}, 10));
    CHECK(q.pop(lits, count, id, st) && id == 10 && count == 1 && lits[0] == lit(1));
    CHECK(q.pop(lits, count, id, st) && id == 11 && count == 2 && lits[1] == lit(3));
    CHECK(!q.pop(lits, count, id, st));
    CHECK(q.last_all_reported() == 1);
    CHECK(st[reportedClauses] == 2 && st[reportedClausesUnit] == 1 &&
          st[reportedClausesBinary] == 1);
    // assignments 1 and 2 were sent before the batch for 1 was read: clause 10
    // reported again for 2 is handed out once (same clause, same solver)
    q.assig_was_sent(1);
    q.assig_was_sent(2);
    q.publish(batch(q, 1, 1, {{lit(1)}}, 10));
    CHECK(q.pop(lits, count, id, st) && id == 10);
    CHECK(!q.pop(lits, count, id, st));
    q.publish(batch(q, 2, 1, {{lit(1)}}, 10));
    CHECK(!q.pop(lits, count, id, st));  // still known from the previous batch
    CHECK(q.last_all_reported() == 3);
    // everything before is released now: a new report is handed out again
    q.publish(batch(q, 3, 1, {{lit(1)}}, 10));
    CHECK(q.pop(lits, count, id, st) && id == 10);
    CHECK(!q.pop(lits, count, id, st));
    // own clause: suppressed until assignment last_sent + 1 is fully reported
    q.assig_was_sent(5);
    q.clause_was_added(77);
    q.publish(batch(q, 4, 2, {{lit(4)}}, 77));  // assignments 4, 5
    CHECK(!q.pop(lits, count, id, st));
}

// Batches are handed out in the order they were published, and one published
// after the earlier ones were drained comes next (ConcurrentQueueTest.cc:29-70:
// iteration and removal both follow insertion order).
static void report_queue_fifo() {
    std::atomic<unsigned long> st[6];
    for (auto &s : st) s = 0;
    ReportQueue q;
    int *lits;
    int count;
    long id;
    for (long a = 0; a < 4; a++) q.assig_was_sent(a);
    q.publish(batch(q, 0, 1, {{lit(3)}}, 3));
    q.publish(batch(q, 1, 1, {{lit(2)}}, 2));
    q.publish(batch(q, 2, 1, {{lit(5)}}, 5));
    CHECK(q.pop(lits, count, id, st) && id == 3 && lits[0] == lit(3));
    CHECK(q.pop(lits, count, id, st) && id == 2 && lits[0] == lit(2));
    CHECK(q.pop(lits, count, id, st) && id == 5 && lits[0] == lit(5));
    CHECK(!q.pop(lits, count, id, st));
    q.publish(batch(q, 3, 1, {{lit(9)}}, 9));
    CHECK(q.pop(lits, count, id, st) && id == 9 && lits[0] == lit(9));
    CHECK(!q.pop(lits, count, id, st));
    CHECK(q.last_all_reported() == 4);
    CHECK(st[reportedClauses] == 4);
}

int main() {
    two_assignments();
    two_solvers();
    many_assignments(64);
    many_assignments(32);
    carry_forward();
    activity_basics();
    approx_nth();
    act_only_threshold();
    act_only_huge_differences();
    rescale();
    reduce_keeps_active();
    report_queue();
    report_queue_fifo();
    std::printf("%d checks, %d failed\n", checks, failures);
    return failures ? 1 : 0;
}
