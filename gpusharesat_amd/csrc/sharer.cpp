// GpuShare::GpuClauseSharer over the gss check engine (see sharer_impl.h for
// the structure and the reference files each part follows).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <thread>

#include "sharer_impl.h"

namespace GpuShare {

namespace {

constexpr float kRescale = 1e19f;  // RESCALE_CONST, gpuShareLib/Clauses.cuh:39
constexpr int kMaxSolvers = GSS_MAX_WORDS;

long now_micros() {
    return std::chrono::duration_cast<std::chrono::microseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

struct Gauge {  // adds elapsed microseconds to a stat (Profiler.h TimeGauge)
    std::atomic<unsigned long> &stat;
    bool on;
    long t0;
    Gauge(std::atomic<unsigned long> &s, bool enabled) : stat(s), on(enabled), t0(now_micros()) {}
    ~Gauge() {
        if (on) stat += (unsigned long)(now_micros() - t0);
    }
};

const char *kGlobalNames[] = {
#define X(v) #v,
#include "gpuShareLib/GlobalStats.h"
#undef X
};
const char *kSolverNames[] = {
#define X(v) #v,
#include "gpuShareLib/OneSolverStats.h"
#undef X
};
constexpr int kGlobalCount = sizeof(kGlobalNames) / sizeof(kGlobalNames[0]);
constexpr int kSolverCount = sizeof(kSolverNames) / sizeof(kSolverNames[0]);

void check_rc(int rc, const char *what) {
    if (rc != GSS_OK)
        throw std::runtime_error(std::string(what) + ": " + gss_last_error());
}

}  // namespace

SharerLimits SharerLimits::from_env() {
    SharerLimits l;
    const char *c = std::getenv("GPUSHARESAT_COMPAT");
    if (c && *c && std::strcmp(c, "0") != 0) {
        l.slots = 32;
        l.max_clause_size = 100;
    }
    return l;
}

/* ---- SolverWindow (gpuShareLib/Assigs.cu:146-300) ------------------------ */

SolverWindow::SolverWindow(int slots)
    : slots_(slots),
      slot_bits_(slots >= 64 ? ~0ull : ((1ull << slots) - 1)),
      not_completed_(~0ull) {}

void SolverWindow::set_var_count(int n) {
    if (n > (int)last_val_.size()) {
        last_val_.resize(n, kValUndef);
        update_pos_.resize(n, -1);
    }
}

// A new update for a variable gives the slots not yet completed the new value
// and the completed slots of this window the value they were sent with, which
// is the variable's last value since it has no update yet in this window; an
// existing update only changes the slots not yet completed (Assigs.cu:163-185).
void SolverWindow::set_var(int var, uint8_t val) {
    if (var >= (int)last_val_.size()) set_var_count(var + 1);
    const uint64_t nc = not_completed_;
    const uint64_t def_m = val != kValUndef ? ~0ull : 0ull;
    const uint64_t true_m = val == kValTrue ? ~0ull : 0ull;
    const int pos = update_pos_[var];
    if (pos < 0 || pos >= (int)updates_.size() || updates_[pos].var != (uint32_t)var) {
        const uint8_t old = last_val_[var];
        gss_var_update u;
        u.var = (uint32_t)var;
        u.word = 0;
        u.is_def = (nc & def_m) | (~nc & (old != kValUndef ? ~0ull : 0ull));
        u.is_true = (nc & true_m) | (~nc & (old == kValTrue ? ~0ull : 0ull));
        update_pos_[var] = (int)updates_.size();
        updates_.push_back(u);
    } else {
        gss_var_update &u = updates_[pos];
        u.is_def = (u.is_def & ~nc) | (nc & def_m);
        u.is_true = (u.is_true & ~nc) | (nc & true_m);
    }
    last_val_[var] = val;
}

long SolverWindow::done() {
    const int pos = (int)(current_id_ % slots_);
    not_completed_ &= ~(1ull << pos);
    return current_id_++;
}

void SolverWindow::current_assignment(uint8_t *out) const {
    if (!last_val_.empty()) std::memcpy(out, last_val_.data(), last_val_.size());
}

SolverWindow::Taken SolverWindow::take(std::vector<gss_var_update> &out, int word) {
    Taken t;
    t.start_id = first_id_;
    t.count = current_id_ - first_id_;
    for (gss_var_update u : updates_) {
        u.word = (uint32_t)word;
        out.push_back(u);
    }
    // slot whose values every slot takes after the check (Assigs.cu:264-271)
    long last;
    if (current_id_ == 0) last = 0;
    else if (current_id_ == first_id_ + slots_) last = current_id_ - 1;
    else last = current_id_;
    t.valid = ~not_completed_ & slot_bits_;
    t.last_pos = updates_.empty() ? -1 : (int)(last % slots_);
    updates_.clear();
    first_id_ = current_id_;
    not_completed_ = ~0ull;
    return t;
}

/* ---- ClauseMirror (gpuShareLib/Clauses.cu) -------------------------------- */

long ClauseMirror::stage(const int *lits, int count, int lbd) {
    std::lock_guard<std::mutex> g(stage_lock_);
    const long id = next_id_++;
    staged_lits_.insert(staged_lits_.end(), lits, lits + count);
    staged_lens_.push_back(count);
    staged_lbd_.push_back(lbd);
    staged_ids_.push_back(id);
    return id;
}

bool ClauseMirror::has_staged() {
    std::lock_guard<std::mutex> g(stage_lock_);
    return !staged_lens_.empty();
}

void ClauseMirror::commit_staged(Batch &b) {
    std::vector<int32_t> lits;
    std::vector<int> lens, lbds;
    std::vector<long> ids;
    {
        std::lock_guard<std::mutex> g(stage_lock_);
        lits.swap(staged_lits_);
        lens.swap(staged_lens_);
        lbds.swap(staged_lbd_);
        ids.swap(staged_ids_);
    }
    b.lits.clear();
    b.lens.clear();
    b.rows.clear();
    b.max_lit = -1;
    size_t p = 0;
    for (size_t i = 0; i < lens.size(); i++) {
        const int len = lens[i];
        // activity decays once per added clause, and a new clause starts at
        // the current increment (Clauses.cu:204,334)
        act_incr_ /= decay_;
        if ((int)lits_.size() <= len) {
            lits_.resize(len + 1);
            meta_.resize(len + 1);
        }
        const uint32_t row = (uint32_t)meta_[len].size();
        meta_[len].push_back(ClauseMeta{ids[i], lbds[i], act_incr_});
        for (int k = 0; k < len; k++) {
            const uint32_t l = (uint32_t)lits[p + k];
            lits_[len].push_back(l);
            b.lits.push_back(l);
            b.max_lit = std::max<int64_t>(b.max_lit, l);
        }
        p += len;
        b.lens.push_back(len);
        b.rows.push_back(row);
        count_++;
        len_sum_ += len;
        added_++;
        if (act_incr_ > kRescale) rescale();
    }
}

void ClauseMirror::bump(int len, uint32_t row) {
    float &a = meta_[len][row].activity;
    a += act_incr_;
    if (a > kRescale) rescale();
}

void ClauseMirror::rescale() {
    for (auto &bucket : meta_)
        for (ClauseMeta &m : bucket) m.activity /= kRescale;
    act_incr_ /= kRescale;
}

// Approximate activity of the n-th least active clause among lbds in
// [min_lbd, max_lbd): a 20000-bucket histogram over log(activity) between
// log(FLT_MIN) and log(FLT_MAX), rounded up to the bucket's upper edge
// (Clauses.cu:484-525).
float ClauseMirror::approx_nth_activity(int min_lbd, int max_lbd, long n) const {
    if (n == 0) return 0.0f;
    const int buckets = 20000;
    std::vector<long> counts(buckets, 0);
    const float lowest = std::log(std::numeric_limits<float>::min());
    const float largest = std::log(std::numeric_limits<float>::max());
    const float step = (largest - lowest) / buckets;
    for (size_t len = 1; len < meta_.size(); len++) {
        for (const ClauseMeta &m : meta_[len]) {
            if (m.lbd < min_lbd || m.lbd >= max_lbd) continue;
            const double v = std::log(m.activity);
            int b = (int)std::floor((v - (double)lowest) / (double)step);
            if (b < 0) b = 0;
            if (b >= buckets) b = buckets - 1;
            counts[b]++;
        }
    }
    long seen = 0;
    for (int b = 0; b < buckets; b++) {
        seen += counts[b];
        if (seen >= n) return std::exp(lowest + (float)(b + 1) * step);
    }
    throw std::logic_error("approx_nth_activity: fewer clauses than requested");
}

// Keeps, in buckets of length >= 3, the clauses with lbd < min_lbd or with
// lbd < max_lbd and activity at least the approximate median; buckets 1 and 2
// are never touched (Clauses.cu:249-282,391-398,426-465 with actOnly).
std::vector<int> ClauseMirror::reduce(int max_lbd) {
    std::vector<int> changed;
    const int min_lbd = 0;
    const float act = approx_nth_activity(min_lbd, max_lbd, count_ / 2);
    for (int len = (int)meta_.size() - 1; len >= 3; len--) {
        std::vector<ClauseMeta> &meta = meta_[len];
        std::vector<uint32_t> &lits = lits_[len];
        size_t to = 0;
        for (size_t r = 0; r < meta.size(); r++) {
            const ClauseMeta &m = meta[r];
            if (m.lbd < min_lbd || (m.lbd < max_lbd && m.activity >= act)) {
                if (to != r) {
                    meta[to] = m;
                    std::memmove(&lits[to * len], &lits[r * len], len * sizeof(uint32_t));
                }
                to++;
            }
        }
        if (to != meta.size()) {
            const long removed = (long)(meta.size() - to);
            count_ -= removed;
            len_sum_ -= removed * len;
            meta.resize(to);
            lits.resize(to * len);
            changed.push_back(len);
        }
    }
    return changed;
}

/* ---- ReportQueue (gpuShareLib/Reported.cu:97-204) ------------------------- */

ClauseBatch *ReportQueue::get_new() {
    std::lock_guard<std::mutex> g(m_);
    ClauseBatch *b;
    if (!pool_.empty()) {
        b = pool_.back();
        pool_.pop_back();
    } else {
        b = new ClauseBatch();
    }
    b->clear();
    return b;
}

void ReportQueue::publish(ClauseBatch *b) {
    std::lock_guard<std::mutex> g(m_);
    ready_.push_back(b);
}

void ReportQueue::clause_was_added(long id) {
    dont_import_.push({id, last_sent_.load() + 1});
    no_import_.insert(id);
}

bool ReportQueue::pop(int *&lits, int &count, long &id, std::atomic<unsigned long> *stats) {
    for (;;) {
        if (!current_) {
            std::lock_guard<std::mutex> g(m_);
            if (ready_.empty()) return false;
            current_ = ready_.front();
            ready_.pop_front();
            taken_.push_back(current_);
        }
        ClauseBatch &b = *current_;
        while (b.next < b.clauses.size()) {
            const auto &c = b.clauses[b.next];
            const int start = c.second;
            const int end = b.next + 1 < b.clauses.size() ? b.clauses[b.next + 1].second
                                                          : (int)b.lits.size();
            b.next++;
            // a clause reported for several assignments (of this or of an
            // earlier batch) is handed out once
            if (no_import_.insert(c.first).second) {
                lits = b.lits.data() + start;
                count = end - start;
                id = c.first;
                stats[reportedClauses]++;
                if (count == 1) stats[reportedClausesUnit]++;
                if (count == 2) stats[reportedClausesBinary]++;
                return true;
            }
        }
        // Batch consumed.  Assignments sent from now on know its clauses; once
        // every assignment that did not is fully reported, its clauses may be
        // reported again (only if the solver deleted them).
        b.knows_from = last_sent_.load() + 1;
        const long seen_all_until = b.start_assig + b.assig_count;
        while (!taken_.empty() && taken_.front()->knows_from <= seen_all_until) {
            ClauseBatch *old = taken_.front();
            for (const auto &c : old->clauses) no_import_.erase(c.first);
            taken_.pop_front();
            std::lock_guard<std::mutex> g(m_);
            pool_.push_back(old);
        }
        while (!dont_import_.empty() && dont_import_.front().second < seen_all_until) {
            no_import_.erase(dont_import_.front().first);
            dont_import_.pop();
        }
        last_all_reported_ = seen_all_until;
        current_ = nullptr;
    }
}

ReportQueue::~ReportQueue() {
    std::lock_guard<std::mutex> g(m_);
    std::set<ClauseBatch *> all(pool_.begin(), pool_.end());
    all.insert(ready_.begin(), ready_.end());
    all.insert(taken_.begin(), taken_.end());
    if (current_) all.insert(current_);
    for (ClauseBatch *b : all) delete b;
}

/* ---- EngineClauseSharer ---------------------------------------------------- */

EngineClauseSharer::EngineClauseSharer(GpuClauseSharerOptions opts,
                                       std::function<void(const std::string &)> log_func)
    : opts_(opts),
      log_func_(std::move(log_func)),
      limits_(SharerLimits::from_env()),
      clauses_(0.99999f),
      global_stats_(kGlobalCount) {
    // option defaults and validation (GpuClauseSharerImpl.cu:40-65)
    if (opts_.minGpuLatencyMicros < 0) opts_.minGpuLatencyMicros = 50;
    if (opts_.clauseActivityDecay < 0) opts_.clauseActivityDecay = 0.99999;
    if (opts_.clauseActivityDecay >= 1)
        throw std::invalid_argument("Clause activity decay must be strictly smaller than 1");
    if (opts_.initReportCountPerCategory < 0) opts_.initReportCountPerCategory = 10;
    if (opts_.initReportCountPerCategory == 0)
        throw std::invalid_argument("initReportCountPerCategory must not be 0");
    if (opts_.gpuThreadsPerBlockGuideline == 0)
        throw std::invalid_argument("gpuThreadsPerBlockGuideline must not be 0");
    if (opts_.gpuBlockCountGuideline == 0)
        throw std::invalid_argument("gpuBlockCountGuideline must not be 0");
    clauses_.set_decay((float)opts_.clauseActivityDecay);
    for (auto &s : global_stats_) s = 0;
    check_rc(gss_engine_create(0, &engine_), "gss_engine_create");
    if (opts_.gpuBlockCountGuideline > 0 && opts_.gpuThreadsPerBlockGuideline > 0) {
        // waves per CU from the guidelines (the persistent grid is sized on
        // the device's CU count)
        int cu = 1;
        gss_engine_cu_count(engine_, &cu);
        const long waves = (long)opts_.gpuBlockCountGuideline *
                           std::max(1, opts_.gpuThreadsPerBlockGuideline / 64);
        gss_set_tuning(engine_, GSS_TUNE_WAVES_PER_CU,
                       (int)std::max<long>(4, std::min<long>(64, waves / std::max(1, cu))));
    }
    if (opts_.maxPageLockedMemory < 0) {
        size_t f = 0, t = 0;
        gss_engine_mem_info(engine_, &f, &t);
        opts_.maxPageLockedMemory = (int)std::min<size_t>(t / 3, (size_t)INT32_MAX);
    }
    log(1, "c GPU clause sharing on MI355X engine, " + std::to_string(limits_.slots) +
               " assignment slots per solver");
    setCpuSolverCount(1);
}

EngineClauseSharer::~EngineClauseSharer() {
    if (engine_) {
        gss_engine_sync(engine_);
        gss_engine_destroy(engine_);
    }
}

void EngineClauseSharer::log(int level, const std::string &msg) {
    if (opts_.verbosity >= level && log_func_) log_func_(msg);
}

int EngineClauseSharer::words_for(int solvers) const {
    int w = 1;
    while (w < solvers) w *= 2;
    return w;
}

void EngineClauseSharer::ensure_table() {
    const int words = words_for(solver_count());
    int64_t vars = std::max<int64_t>(var_count_, table_vars_);
    for (auto &w : windows_) vars = std::max<int64_t>(vars, w->var_count());
    vars = std::max<int64_t>(vars, clause_vars_);
    if (words != table_words_) {
        if (table_words_ == 0) {
            check_rc(gss_assig_reset(engine_, vars, words), "gss_assig_reset");
        } else {
            // re-layout: keep each existing solver's word
            std::vector<uint64_t> od((size_t)table_vars_ * table_words_),
                ot((size_t)table_vars_ * table_words_);
            if (table_vars_)
                check_rc(gss_assig_download(engine_, od.data(), ot.data()), "gss_assig_download");
            std::vector<uint64_t> nd((size_t)vars * words, 0), nt((size_t)vars * words, 0);
            const int keep = std::min(words, table_words_);
            for (int64_t v = 0; v < table_vars_; v++)
                for (int w = 0; w < keep; w++) {
                    nd[v * words + w] = od[v * table_words_ + w];
                    nt[v * words + w] = ot[v * table_words_ + w];
                }
            check_rc(gss_assig_reset(engine_, vars, words), "gss_assig_reset");
            if (vars) check_rc(gss_assig_upload(engine_, nd.data(), nt.data()), "gss_assig_upload");
        }
        table_words_ = words;
        table_vars_ = vars;
    } else if (vars > table_vars_) {
        check_rc(gss_assig_grow_vars(engine_, vars), "gss_assig_grow_vars");
        table_vars_ = vars;
    }
}

bool EngineClauseSharer::start_run(InFlight &run) {
    ClauseMirror::Batch b;
    clauses_.commit_staged(b);
    global_stats_[gpuClausesAdded] = (unsigned long)clauses_.added();
    global_stats_[gpuClauses] = (unsigned long)clauses_.clause_count();
    global_stats_[gpuClauseLengthSum] = (unsigned long)clauses_.length_sum();
    if (!b.lens.empty()) {
        clause_vars_ = std::max<int64_t>(clause_vars_, (b.max_lit >> 1) + 1);
        std::vector<uint32_t> rows(b.lens.size());
        int rc = gss_db_append(engine_, (int64_t)b.lens.size(), b.lits.data(), b.lens.data(),
                               rows.data());
        if (rc == GSS_E_NOMEM) {
            // out of device memory: halve the DB and re-upload (GpuRunner.cu:243-260)
            ran_out_of_memory_ = true;
            reduce_db_now();
            return false;
        }
        check_rc(rc, "gss_db_append");
        for (size_t i = 0; i < rows.size(); i++)
            if (rows[i] != b.rows[i])
                throw std::logic_error("device clause rows diverged from the host mirror");
    }
    if (clauses_.clause_count() == 0) return false;  // GpuRunner.cu:281-286
    ensure_table();
    const int S = solver_count();
    run.taken.assign(S, SolverWindow::Taken());
    run.updates.clear();
    run.last_pos.assign(table_words_, -1);
    std::vector<uint64_t> valid(table_words_, 0);
    {
        Gauge g(global_stats_[timeSpentFillingAssigs], opts_.quickProf);
        for (int s = 0; s < S; s++) {
            SolverWindow &w = *windows_[s];
            // a solver busy setting values waits for the next run (Assigs.cu:350)
            if (!w.lock.try_lock()) continue;
            run.taken[s] = w.take(run.updates, s);
            w.lock.unlock();
            valid[s] = run.taken[s].valid;
            run.last_pos[s] = (int8_t)run.taken[s].last_pos;
        }
    }
    // table vars may have grown through the windows while we held no lock
    ensure_table();
    check_rc(gss_assig_apply(engine_, (int64_t)run.updates.size(), run.updates.data()),
             "gss_assig_apply");
    check_rc(gss_sweep_async(engine_, valid.data()), "gss_sweep_async");
    check_rc(gss_assig_roll_forward_applied(engine_, run.last_pos.data()),
             "gss_assig_roll_forward_applied");
    run.clause_count = clauses_.clause_count();
    run.active = true;
    return true;
}

void EngineClauseSharer::collect(InFlight &run, std::vector<gss_hit> &hits) {
    (void)run;
    int64_t n = 0;
    check_rc(gss_sweep_wait(engine_, &n), "gss_sweep_wait");
    hits.resize((size_t)n);
    if (n) check_rc(gss_hits_copy(engine_, hits.data(), n, &n), "gss_hits_copy");
    if (opts_.quickProf) {
        float ms = 0;
        gss_sweep_kernel_ms(engine_, &ms);
        global_stats_[timeSpentTestingClauses] += (unsigned long)(ms * 1000.0f);
    }
}

void EngineClauseSharer::deliver(InFlight &run, std::vector<gss_hit> &hits) {
    global_stats_[gpuRuns]++;
    long assigs = 0, solvers_with = 0;
    for (const auto &t : run.taken) {
        assigs += t.count;
        if (t.count) solvers_with++;
    }
    const unsigned long cl = (unsigned long)run.clause_count;
    global_stats_[totalAssigClauseTested] += cl * (unsigned long)assigs;
    global_stats_[clauseTestsOnGroups] += cl;
    global_stats_[clauseTestsOnAssigs] += cl * (unsigned long)solvers_with;
    global_stats_[gpuReports] += hits.size();
    for (const gss_hit &h : hits) clauses_.bump(h.len, h.row);
    Gauge g(global_stats_[timeSpentFillingReported], opts_.quickProf);
    const int S = (int)run.taken.size();
    std::vector<ClauseBatch *> batch(S, nullptr);
    for (int s = 0; s < S; s++) {
        if (run.taken[s].count <= 0) continue;
        batch[s] = reports_[s]->get_new();
        batch[s]->start_assig = run.taken[s].start_id;
        batch[s]->assig_count = run.taken[s].count;
    }
    for (const gss_hit &h : hits) {
        const int s = h.word;
        if (s >= S || !batch[s]) continue;  // cannot happen: hits only in valid slots
        ClauseBatch &b = *batch[s];
        b.clauses.push_back({clauses_.meta(h.len, h.row).id, (int)b.lits.size()});
        const uint32_t *l = clauses_.row_lits(h.len, h.row);
        b.lits.insert(b.lits.end(), l, l + h.len);
    }
    for (int s = 0; s < S; s++)
        if (batch[s]) reports_[s]->publish(batch[s]);
}

void EngineClauseSharer::run_once(bool can_start) {
    InFlight prev;
    const bool had = in_flight_.active;
    if (had) {
        prev = std::move(in_flight_);
        in_flight_ = InFlight();
        collect(prev, hit_buf_);
    }
    if (can_start) start_run(in_flight_);
    if (had) deliver(prev, hit_buf_);
}

void EngineClauseSharer::gpuRun() {
    const long t0 = now_micros();
    run_once(true);
    const long spent = now_micros() - t0;
    if (spent < opts_.minGpuLatencyMicros)
        std::this_thread::sleep_for(std::chrono::microseconds(opts_.minGpuLatencyMicros - spent));
}

void EngineClauseSharer::reduce_db_now() {
    Gauge g(global_stats_[timeSpentReduceGpuDb], true);
    added_at_last_reduce_ = clauses_.added();
    global_stats_[gpuReduceDbs]++;
    log(2, "c Reducing gpu clause db, clause count is " + std::to_string(clauses_.clause_count()));
    for (int len : clauses_.reduce(limits_.max_clause_size)) {
        const auto &lits = clauses_.bucket_lits(len);
        check_rc(gss_db_set_bucket(engine_, len, (int64_t)(lits.size() / len), lits.data()),
                 "gss_db_set_bucket");
    }
    global_stats_[gpuClauses] = (unsigned long)clauses_.clause_count();
    global_stats_[gpuClauseLengthSum] = (unsigned long)clauses_.length_sum();
    log(2, "c Done reducing gpu clause db, clause count is " +
               std::to_string(clauses_.clause_count()));
}

void EngineClauseSharer::reduceDb() {
    // hits of a run in flight name clauses by position: deliver them first
    // (GpuClauseSharerImpl.cu:105-110)
    run_once(false);
    reduce_db_now();
}

long EngineClauseSharer::getAddedClauseCount() { return clauses_.added(); }

long EngineClauseSharer::getAddedClauseCountAtLastReduceDb() { return added_at_last_reduce_; }

bool EngineClauseSharer::hasRunOutOfGpuMemoryOnce() { return ran_out_of_memory_; }

void EngineClauseSharer::getGpuMemInfo(size_t &free, size_t &total) {
    check_rc(gss_engine_mem_info(engine_, &free, &total), "gss_engine_mem_info");
}

int EngineClauseSharer::getGlobalStatCount() { return kGlobalCount; }

long EngineClauseSharer::getGlobalStat(GlobalStats stat) {
    if ((int)stat < 0 || (int)stat >= kGlobalCount) return 0;
    return (long)global_stats_[stat].load();
}

const char *EngineClauseSharer::getGlobalStatName(GlobalStats stat) {
    if ((int)stat < 0 || (int)stat >= kGlobalCount) return nullptr;
    return kGlobalNames[stat];
}

int EngineClauseSharer::getOneSolverStatCount() { return kSolverCount; }

const char *EngineClauseSharer::getOneSolverStatName(OneSolverStats stat) {
    if ((int)stat < 0 || (int)stat >= kSolverCount) return nullptr;
    return kSolverNames[stat];
}

long EngineClauseSharer::getOneSolverStat(int solver, OneSolverStats stat) {
    if ((int)stat < 0 || (int)stat >= kSolverCount) return 0;
    return (long)solver_stats_[solver][stat].load();
}

// The header line goes to `file`, followed by one DIMACS line per clause,
// bucket by bucket (Clauses.cu:527-547 prints the header to stdout instead).
void EngineClauseSharer::writeClausesInCnf(FILE *file) {
    if (!file) return;
    // staged clauses are not on the device yet; they are written once committed
    std::fprintf(file, "p cnf %d %ld\n", var_count_, clauses_.clause_count());
    for (int len = 1; len <= clauses_.max_len(); len++) {
        const int64_t n = clauses_.rows(len);
        for (int64_t r = 0; r < n; r++) {
            const uint32_t *l = clauses_.row_lits(len, (uint32_t)r);
            for (int k = 0; k < len; k++)
                std::fprintf(file, "%d ", (int)((l[k] >> 1) + 1) * ((l[k] & 1) ? -1 : 1));
            std::fprintf(file, "0\n");
        }
    }
}

void EngineClauseSharer::setVarCount(int n) {
    var_count_ = n;
    for (auto &w : windows_) {
        std::lock_guard<std::mutex> g(w->lock);
        w->set_var_count(n);
    }
}

long EngineClauseSharer::addClause(int solverId, int *lits, int count) {
    if (count < 1 || count > limits_.max_clause_size || !lits) return -1;
    for (int i = 0; i < count; i++)
        if (lits[i] < 0) return -1;
    const long id = clauses_.stage(lits, count, count);  // lbd = size (GpuClauseSharerImpl.cu:125)
    if (solverId >= 0 && solverId < solver_count()) reports_[solverId]->clause_was_added(id);
    return id;
}

void EngineClauseSharer::setCpuSolverCount(int count) {
    if (count < 1 || count > kMaxSolvers)
        throw std::invalid_argument("solver count must be in [1, " + std::to_string(kMaxSolvers) +
                                    "]");
    if (count < solver_count()) throw std::invalid_argument("solver count can only grow");
    run_once(false);  // the table width may change: nothing may be in flight
    const int old = solver_count();
    for (int s = old; s < count; s++) {
        windows_.push_back(std::make_unique<SolverWindow>(limits_.slots));
        windows_.back()->set_var_count(var_count_);
        reports_.push_back(std::make_unique<ReportQueue>());
        to_unset_.emplace_back();
        solver_stats_.emplace_back(new std::atomic<unsigned long>[kSolverCount]);
        for (int k = 0; k < kSolverCount; k++) solver_stats_.back()[k] = 0;
    }
}

void EngineClauseSharer::unset_pending_locked(int solver) {
    std::vector<int> &un = to_unset_[solver];
    SolverWindow &w = *windows_[solver];
    for (int l : un) w.set_var(l >> 1, kValUndef);
    solver_stats_[solver][varUpdatesSentToGpu] += un.size();
    un.clear();
}

bool EngineClauseSharer::trySetSolverValues(int solver, int *lits, int count) {
    SolverWindow &w = *windows_[solver];
    std::lock_guard<std::mutex> g(w.lock);
    if (!w.available()) {
        solver_stats_[solver][failuresToFindAssig]++;
        return false;
    }
    unset_pending_locked(solver);
    for (int i = 0; i < count; i++) w.set_var(lits[i] >> 1, (lits[i] & 1) ? kValFalse : kValTrue);
    solver_stats_[solver][varUpdatesSentToGpu] += (unsigned long)count;
    return true;
}

void EngineClauseSharer::unsetSolverValues(int solver, int *lits, int count) {
    SolverWindow &w = *windows_[solver];
    std::lock_guard<std::mutex> g(w.lock);
    if (w.available()) {
        unset_pending_locked(solver);
        for (int i = 0; i < count; i++) w.set_var(lits[i] >> 1, kValUndef);
        solver_stats_[solver][varUpdatesSentToGpu] += (unsigned long)count;
    } else {
        // every slot is taken: apply once one frees (GpuClauseSharerImpl.cu:169-174)
        to_unset_[solver].insert(to_unset_[solver].end(), lits, lits + count);
    }
}

long EngineClauseSharer::trySendAssignment(int solver) {
    SolverWindow &w = *windows_[solver];
    std::lock_guard<std::mutex> g(w.lock);
    if (!w.available()) {
        solver_stats_[solver][failuresToFindAssig]++;
        return -1;
    }
    const long id = w.done();
    solver_stats_[solver][assigsSentToGpu]++;
    reports_[solver]->assig_was_sent(id);
    return id;
}

bool EngineClauseSharer::popReportedClause(int solver, int *&lits, int &count, long &gpuClauseId) {
    return reports_[solver]->pop(lits, count, gpuClauseId, solver_stats_[solver].get());
}

long EngineClauseSharer::getLastAssigAllReported(int solver) {
    return reports_[solver]->last_all_reported();
}

void EngineClauseSharer::getCurrentAssignment(int solver, uint8_t *assig) {
    SolverWindow &w = *windows_[solver];
    {
        std::lock_guard<std::mutex> g(w.lock);
        w.current_assignment(assig);
    }
    for (int l : to_unset_[solver]) assig[l >> 1] = kValUndef;
}

GpuClauseSharer *makeGpuClauseSharerPtr(GpuClauseSharerOptions opts,
                                        std::function<void(const std::string &str)> logFunc) {
    return new EngineClauseSharer(opts, std::move(logFunc));
}

}  // namespace GpuShare
