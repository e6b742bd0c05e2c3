/*
 * CPU oracle for the clause-vs-assignment check.  TEST INFRASTRUCTURE ONLY:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as the checker.  The product path never calls it.
 *
 * What it restates (reference files, relative to the GpuShareSat tree):
 *
 *   orc_rule_clause   gpuShareLib/GpuRunner.cu:34-57 (ReportComputer) and
 *                     GpuRunner.cu:68-89 (dCheckOneClauseOneSolver): per literal
 *                       isFalse = (sign ? isTrue : ~isTrue) & isDef
 *                       jou     = (allFalse & ~isDef) | (jou & isFalse)
 *                       allFalse&= isFalse
 *                     starting from allFalse = valid slots, jou = 0; report
 *                     allFalse | jou.  Literal encoding 2*var + sign, sign 1 =
 *                     negative (gpuShareLib/SolverTypes.h:56-60,
 *                     BaseTypes.cuh:167-173).  Assignment planes are the
 *                     reference's MultiLBool{isDef,isTrue} (BaseTypes.cuh:65-101)
 *                     widened to 64 assignments per uint64 word.
 *   orc_sweep         GpuRunner.cu:134-177 (dFindClauses): every clause against
 *                     every valid slot of every word; emits one record per
 *                     (clause, word) with a nonzero report, sorted by
 *                     (clause, word).  Unlike gpuShareLib/Reporter.cuh:46-48
 *                     nothing is ever dropped.
 *   orc_drand/irand   glucose-syrup/utils/Utils.h:29-43 (Glucose's LCG).
 *   orc_perftest_counts
 *                     glucose-syrup/perftest/perfTest.cu:48-125,153-207: 2M
 *                     clauses of irand(12,20) random literals over 500 vars
 *                     (seed 0.4), then per round a fresh assignment with
 *                     irand(seed,3) per var (0 -> true, 1 -> false, 2 -> undef,
 *                     seed 0.6), one assignment per run.  Returns the cumulative
 *                     report count after each round (the reference asserts 143
 *                     after 15 rounds and 19739 after 2000).
 *
 * Pinning (tests/test_oracle.py): the golden perfTest counts, the reference's
 * known-answer tests (glucose-syrup/test/GpuSolverTest.cu:343-454,678-724),
 * and brute force over every 3-valued assignment of small clauses (duplicates
 * and complementary literals included) against the counting definition of
 * "falsified" / "exactly one undefined, rest false".
 */
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    uint32_t clause;
    uint32_t word;
    uint64_t false_mask;
    uint64_t unit_mask;
} orc_hit;

/* ReportComputer over one clause, one word.  isDef/isTrue are var-major:
 * plane[var * W + w]. */
void orc_rule_clause(const uint32_t *lits, int len, const uint64_t *isDef,
                     const uint64_t *isTrue, int W, int w, uint64_t start,
                     uint64_t *false_mask, uint64_t *unit_mask) {
    uint64_t allFalse = start, jou = 0;
    for (int i = 0; i < len; i++) {
        uint32_t lit = lits[i];
        uint32_t v = lit >> 1;
        uint64_t def = isDef[(size_t)v * W + w];
        uint64_t tru = isTrue[(size_t)v * W + w];
        uint64_t isFalse = ((lit & 1) ? tru : ~tru) & def;
        jou = (allFalse & ~def) | (jou & isFalse);
        allFalse &= isFalse;
        if (!(allFalse | jou)) break; /* nothing left to report (GpuRunner.cu:80-82) */
    }
    *false_mask = allFalse;
    *unit_mask = jou;
}

typedef struct {
    const int64_t *offsets;
    const uint32_t *lits;
    const uint64_t *isDef, *isTrue, *valid;
    int W;
    int64_t c0, c1;
    orc_hit *out;
    int64_t n, cap;
} sweep_job;

static void *sweep_worker(void *arg) {
    sweep_job *j = (sweep_job *)arg;
    j->n = 0;
    for (int64_t c = j->c0; c < j->c1; c++) {
        const uint32_t *l = j->lits + j->offsets[c];
        int len = (int)(j->offsets[c + 1] - j->offsets[c]);
        for (int w = 0; w < j->W; w++) {
            if (!j->valid[w]) continue;
            uint64_t fm, um;
            orc_rule_clause(l, len, j->isDef, j->isTrue, j->W, w, j->valid[w], &fm, &um);
            if (fm | um) {
                if (j->n == j->cap) {
                    int64_t nc = j->cap ? j->cap * 2 : 4096;
                    orc_hit *no = (orc_hit *)realloc(j->out, (size_t)nc * sizeof(orc_hit));
                    if (!no) return (void *)1;
                    j->out = no;
                    j->cap = nc;
                }
                orc_hit *h = &j->out[j->n++];
                h->clause = (uint32_t)c;
                h->word = (uint32_t)w;
                h->false_mask = fm;
                h->unit_mask = um;
            }
        }
    }
    return NULL;
}

/* Sweep all clauses (CSR: offsets[nclauses+1] into lits) against W words.
 * Writes up to cap records to out, sorted by (clause, word); returns the total
 * number of records (call again with a larger buffer when it exceeds cap), or
 * -1 on allocation failure. */
int64_t orc_sweep(int64_t nclauses, const int64_t *offsets, const uint32_t *lits,
                  int W, const uint64_t *isDef, const uint64_t *isTrue,
                  const uint64_t *valid, int nthreads, orc_hit *out, int64_t cap) {
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    if ((int64_t)nthreads > nclauses) nthreads = nclauses > 0 ? (int)nclauses : 1;
    sweep_job jobs[256];
    pthread_t th[256];
    for (int t = 0; t < nthreads; t++) {
        sweep_job *j = &jobs[t];
        memset(j, 0, sizeof(*j));
        j->offsets = offsets;
        j->lits = lits;
        j->isDef = isDef;
        j->isTrue = isTrue;
        j->valid = valid;
        j->W = W;
        j->c0 = nclauses * t / nthreads;
        j->c1 = nclauses * (t + 1) / nthreads;
    }
    int failed = 0;
    for (int t = 1; t < nthreads; t++)
        if (pthread_create(&th[t], NULL, sweep_worker, &jobs[t])) {
            /* run it inline instead */
            th[t] = 0;
            if (sweep_worker(&jobs[t])) failed = 1;
        }
    if (sweep_worker(&jobs[0])) failed = 1;
    for (int t = 1; t < nthreads; t++) {
        if (th[t]) {
            void *r;
            pthread_join(th[t], &r);
            if (r) failed = 1;
        }
    }
    int64_t total = 0;
    for (int t = 0; t < nthreads; t++) {
        if (!failed && out) {
            for (int64_t i = 0; i < jobs[t].n; i++)
                if (total + i < cap) out[total + i] = jobs[t].out[i];
        }
        total += jobs[t].n;
        free(jobs[t].out);
    }
    return failed ? -1 : total;
}

/* ---- Glucose LCG, glucose-syrup/utils/Utils.h:29-43 ---------------------- */
double orc_drand(double *seed) {
    *seed *= 1389796;
    int q = (int)(*seed / 2147483647);
    *seed -= (double)q * 2147483647;
    return *seed / 2147483647;
}

int orc_irand(double *seed, int size) { return (int)(orc_drand(seed) * size); }

/* Generates the perfTest clause set: clause sizes irand(seed, minSize, maxSize),
 * literals mkLit(irand(seed, nVars), irand(seed, 2)).  The argument evaluation
 * order of that call is unspecified in C++; sign_first != 0 draws the sign
 * before the variable.  lits must hold nclauses * (maxSize - 1) entries;
 * offsets nclauses + 1. */
void orc_perftest_clauses(int nclauses, int minSize, int maxSize, int nVars,
                          double seed, int sign_first, int64_t *offsets,
                          uint32_t *lits) {
    int64_t pos = 0;
    for (int c = 0; c < nclauses; c++) {
        offsets[c] = pos;
        int size = orc_irand(&seed, maxSize - minSize) + minSize;
        for (int l = 0; l < size; l++) {
            int v, s;
            if (sign_first) {
                s = orc_irand(&seed, 2);
                v = orc_irand(&seed, nVars);
            } else {
                v = orc_irand(&seed, nVars);
                s = orc_irand(&seed, 2);
            }
            lits[pos++] = (uint32_t)(2 * v + (s ? 1 : 0));
        }
    }
    offsets[nclauses] = pos;
}

/* perfTest assignments: for round r, var i: p = irand(seed, 3); p == 0 -> the
 * solver enqueues mkLit(i, false) (var true), p == 1 -> mkLit(i, true) (var
 * false), p == 2 -> unassigned (perfTest.cu:60-73).  Assignment r is slot
 * r % 64 of word r / 64.  isDef/isTrue are [nVars][W] with W = ceil(n/64). */
void orc_perftest_assigs(int nrounds, int nVars, double seed, uint64_t *isDef,
                         uint64_t *isTrue) {
    int W = (nrounds + 63) / 64;
    memset(isDef, 0, sizeof(uint64_t) * (size_t)nVars * W);
    memset(isTrue, 0, sizeof(uint64_t) * (size_t)nVars * W);
    for (int r = 0; r < nrounds; r++) {
        for (int v = 0; v < nVars; v++) {
            int p = orc_irand(&seed, 3);
            uint64_t bit = 1ull << (r % 64);
            if (p == 0 || p == 1) {
                isDef[(size_t)v * W + r / 64] |= bit;
                if (p == 0) isTrue[(size_t)v * W + r / 64] |= bit;
            }
        }
    }
}

/* Cumulative report counts after each perfTest round (counts[nrounds]). */
int orc_perftest_counts(int nclauses, int minSize, int maxSize, int nVars,
                        int nrounds, int sign_first, int nthreads, int64_t *counts) {
    int W = (nrounds + 63) / 64;
    int64_t *offsets = (int64_t *)malloc(sizeof(int64_t) * ((size_t)nclauses + 1));
    uint32_t *lits = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)nclauses * (maxSize - 1));
    uint64_t *isDef = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)nVars * W);
    uint64_t *isTrue = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)nVars * W);
    uint64_t *valid = (uint64_t *)calloc((size_t)W, sizeof(uint64_t));
    if (!offsets || !lits || !isDef || !isTrue || !valid) return -1;
    orc_perftest_clauses(nclauses, minSize, maxSize, nVars, 0.4, sign_first, offsets, lits);
    orc_perftest_assigs(nrounds, nVars, 0.6, isDef, isTrue);
    for (int r = 0; r < nrounds; r++) valid[r / 64] |= 1ull << (r % 64);
    int64_t n = orc_sweep(nclauses, offsets, lits, W, isDef, isTrue, valid, nthreads, NULL, 0);
    orc_hit *hits = n > 0 ? (orc_hit *)malloc(sizeof(orc_hit) * (size_t)n) : NULL;
    int rc = 0;
    if (n < 0 || (n > 0 && !hits)) rc = -1;
    else {
        if (n > 0) orc_sweep(nclauses, offsets, lits, W, isDef, isTrue, valid, nthreads, hits, n);
        int64_t *per = (int64_t *)calloc((size_t)nrounds, sizeof(int64_t));
        for (int64_t i = 0; i < n; i++) {
            uint64_t m = hits[i].false_mask | hits[i].unit_mask;
            while (m) {
                int b = __builtin_ctzll(m);
                m &= m - 1;
                per[hits[i].word * 64 + b]++;
            }
        }
        int64_t acc = 0;
        for (int r = 0; r < nrounds; r++) {
            acc += per[r];
            counts[r] = acc;
        }
        free(per);
    }
    free(hits);
    free(offsets);
    free(lits);
    free(isDef);
    free(isTrue);
    free(valid);
    return rc;
}
