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
    // host copies of at least this size are pageable (CorrespArr.cu:212-235)
    check_rc(gss_engine_set_page_locked_limit(engine_, opts_.maxPageLockedMemory),
             "gss_engine_set_page_locked_limit");
    // reduction policy: the reference's activity median on the host mirror
    // (default), the same rule on the device, or the learned score
    if (const char *p = std::getenv("GPUSHARESAT_REDUCE_POLICY")) {
        const std::string v = p;
        if (v == "device") policy_ = Policy::Device;
        else if (v == "learned") policy_ = Policy::Learned;
        else if (!v.empty() && v != "host")
            throw std::invalid_argument("GPUSHARESAT_REDUCE_POLICY must be host, device or learned");
    }
    // record form of a run's reports: one per clause with its solvers
    // (default) or one per (clause, solver)
    if (const char *r = std::getenv("GPUSHARESAT_RECORDS")) {
        const std::string v = r;
        if (v == "ids") clause_records_ = false;
        else if (!v.empty() && v != "clauses")
            throw std::invalid_argument("GPUSHARESAT_RECORDS must be clauses or ids");
    }
    // runs screened against the table's aggregate first (same reports; it
    // pays where a solver's slots share most of their values)
    if (const char *sc = std::getenv("GPUSHARESAT_SCREEN")) {
        const std::string v = sc;
        if (v == "0") screen_cfg_ = Screen::Off;
        else if (v == "1") screen_cfg_ = Screen::On;
        else if (v.rfind("auto:", 0) == 0 && std::atoi(v.c_str() + 5) >= 8)
            screen_period_ = std::atoi(v.c_str() + 5);
        else if (!v.empty() && v != "auto")
            throw std::invalid_argument("GPUSHARESAT_SCREEN must be 0, 1, auto or auto:N (N >= 8)");
    }
    if (policy_ != Policy::Host) {
        max_tags_ = 1 << 22;
        if (const char *t = std::getenv("GPUSHARESAT_RESIDENCY_MAX_TAGS"))
            max_tags_ = std::max(1ll, std::atoll(t));
        check_rc(gss_db_enable_tags(engine_), "gss_db_enable_tags");
        check_rc(gss_res_create(engine_, max_tags_, (float)opts_.clauseActivityDecay, 1, &res_),
                 "gss_res_create");
        free_tags_.resize((size_t)max_tags_);
        for (int64_t i = 0; i < max_tags_; i++) free_tags_[i] = (uint32_t)(max_tags_ - 1 - i);
        evicted_.assign((size_t)max_tags_, 0);
    }
    log(1, "c GPU clause sharing on MI355X engine, " + std::to_string(limits_.slots) +
               " assignment slots per solver");
    setCpuSolverCount(1);
}

EngineClauseSharer::~EngineClauseSharer() {
    if (engine_) {
        gss_engine_sync(engine_);
        if (res_) gss_res_destroy(res_);  // the scorer goes before its engine
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

// A run (GpuRunner::startGpuRunAsync, GpuRunner.cu:274-351) is one batched
// step: the run's new clauses, the solvers' assignment updates, the valid
// slots and the roll-forward positions are staged into the step's pinned slot
// (one host-to-device copy, made by the step's first kernel), and one HIP
// graph appends, applies, checks, rolls forward and compacts the records.
// The records are gss_hit_id (clause and solver word, no masks: delivery never
// reads them); gss_step_wait starts their copy to pinned host memory on the
// slot's own stream, and deliver() reads them there.
bool EngineClauseSharer::start_run(InFlight &run) {
    ClauseMirror::Batch b;
    clauses_.commit_staged(b);
    global_stats_[gpuClausesAdded] = (unsigned long)clauses_.added();
    global_stats_[gpuClauses] = (unsigned long)clauses_.clause_count();
    global_stats_[gpuClauseLengthSum] = (unsigned long)clauses_.length_sum();
    if (!b.lens.empty()) {
        clause_vars_ = std::max<int64_t>(clause_vars_, (b.max_lit >> 1) + 1);
        if (policy_ != Policy::Host) {
            // device policies: appended now, with tags (gss_res_add)
            if (!add_on_device(b)) return false;
        } else {
            // room for the batch first, before any window is taken: out of
            // device memory, the DB is halved and re-uploaded
            // (GpuRunner.cu:243-260), but only after the run in flight has
            // been delivered: its hits name clauses by row, and the
            // reduction moves rows (GpuRunner.cu:251-253).  The device then
            // holds none of this batch, which the mirror already has.
            const int rc = gss_db_reserve(engine_, (int64_t)b.lens.size(), b.lens.data());
            if (rc == GSS_E_NOMEM) {
                ran_out_of_memory_ = true;
                reduce_pending_ = true;
                dirty_.insert(b.lens.begin(), b.lens.end());
                return false;
            }
            check_rc(rc, "gss_db_reserve");
        }
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
    if (policy_ != Policy::Host) {
        // a plain run: apply, sweep (if some slot is valid); the roll-forward
        // and the records' accounting follow the wait (collect)
        run.plain = true;
        check_rc(gss_assig_apply(engine_, (int64_t)run.updates.size(), run.updates.data()),
                 "gss_assig_apply");
        run.swept = false;
        for (uint64_t v : valid) run.swept = run.swept || v != 0;
        if (run.swept) check_rc(gss_sweep_async(engine_, valid.data()), "gss_sweep_async");
        run.clause_count = clauses_.clause_count();
        run.active = true;
        return true;
    }
    // A run with no clause to append, no update and no valid slot would
    // change nothing on the device (the roll-forward only touches updated
    // variables) and report nothing: it is delivered without a step.
    bool any_valid = false;
    for (uint64_t v : valid) any_valid = any_valid || v != 0;
    if (b.lens.empty() && run.updates.empty() && !any_valid) {
        run.idle = true;
        run.clause_count = clauses_.clause_count();
        run.active = true;
        return true;
    }
    std::vector<uint32_t> rows(b.lens.size());
    gss_step_desc d{};
    d.n_clauses = (int64_t)b.lens.size();
    d.lits = b.lits.data();
    d.lens = b.lens.data();
    d.out_rows = rows.data();
    d.n_ups = (int64_t)run.updates.size();
    d.ups = run.updates.data();
    d.valid = valid.data();
    d.last_pos = run.last_pos.data();
    d.records = clause_records_ ? GSS_RECORDS_CLAUSES : GSS_RECORDS_IDS;
    run.screened = choose_screen();
    d.flags = run.screened ? GSS_STEP_SCREEN : 0;
    check_rc(gss_step_submit(engine_, &d), "gss_step_submit");
    for (size_t i = 0; i < rows.size(); i++)
        if (rows[i] != b.rows[i])
            throw std::logic_error("device clause rows diverged from the host mirror");
    run.clause_count = clauses_.clause_count();
    run.active = true;
    return true;
}

void EngineClauseSharer::collect(InFlight &run) {
    if (run.plain) {
        int64_t n = 0;
        if (run.swept) check_rc(gss_sweep_wait(engine_, &n), "gss_sweep_wait");
        check_rc(gss_assig_roll_forward_applied(engine_, run.last_pos.data()),
                 "gss_assig_roll_forward_applied");
        id_buf_.clear();
        if (run.swept) {
            // one activity bump per record, hit history (GpuRunner.cu:375-378)
            int64_t pairs = 0;
            check_rc(gss_res_account(res_, round_, &pairs), "gss_res_account");
            hit_buf_.resize((size_t)n);
            if (n) check_rc(gss_hits_copy(engine_, hit_buf_.data(), n, &n), "gss_hits_copy");
            id_buf_.resize((size_t)n);
            for (int64_t i = 0; i < n; i++)
                id_buf_[i] = gss_hit_id{hit_buf_[i].row, hit_buf_[i].len, hit_buf_[i].word};
            if (opts_.quickProf) {
                float ms = 0;
                gss_sweep_kernel_ms(engine_, &ms);
                global_stats_[timeSpentTestingClauses] += (unsigned long)(ms * 1000.0f);
            }
        }
        run.records = n;
        if (policy_ == Policy::Learned) train_policy();
        round_++;
        return;
    }
    if (run.idle) {
        run.records = 0;
        return;
    }
    int64_t n = 0;
    float ms = 0.f;
    check_rc(gss_step_wait(engine_, &n, &ms), "gss_step_wait");
    run.records = n;
    if (opts_.quickProf) global_stats_[timeSpentTestingClauses] += (unsigned long)(ms * 1000.0f);
    note_screen_time(run.screened, ms, run.clause_count);
}

bool EngineClauseSharer::choose_screen() {
    if (screen_cfg_ != Screen::Auto) return screen_cfg_ == Screen::On;
    const int64_t k = screen_runs_++;
    if (k > 0 && k % screen_period_ == 0) {
        // a period (and its probe of the other form) ended: the averages of
        // its runs decide, and the next period averages afresh
        const double m[2] = {screen_n_[0] ? screen_sum_[0] / (double)screen_n_[0] : 0.0,
                             screen_n_[1] ? screen_sum_[1] / (double)screen_n_[1] : 0.0};
        const int b = screen_best_ ? 1 : 0;
        if (m[b] > 0.0 && m[1 - b] > 0.0 && m[1 - b] < 0.95 * m[b]) screen_best_ = !screen_best_;
        log(2, "screening: device ns per clause " + std::to_string(m[1] * 1e6) + " screened, " +
                   std::to_string(m[0] * 1e6) + " not; runs " +
                   (screen_best_ ? "screened" : "not screened"));
        screen_sum_[0] = screen_sum_[1] = 0.0;
        screen_n_[0] = screen_n_[1] = 0;
    }
    const bool probe = k % screen_period_ >= screen_period_ - kScreenProbe;
    return probe ? !screen_best_ : screen_best_;
}

void EngineClauseSharer::note_screen_time(bool screened, float ms, long clauses) {
    const int f = screened ? 1 : 0;
    const bool steady = last_screened_ == f;
    last_screened_ = f;
    if (!steady || clauses <= 0 || !(ms > 0.f)) return;
    screen_sum_[f] += (double)ms / (double)clauses;
    screen_n_[f]++;
}

// Learned policy: every 16 runs a snapshot of up to 8192 resident clauses'
// features is taken, and 5 runs later the MLP trains on it (label: reported
// since), as the residency workload does per round.
void EngineClauseSharer::train_policy() {
    if (!failed_lens_.empty()) return;  // the mirror holds clauses the device has not
    const int32_t phase = round_ % 16;
    if (phase == 11 && clauses_.clause_count() > 0) {
        std::vector<uint32_t> tags;
        const long total = clauses_.clause_count();
        const long want = std::min<long>(8192, total);
        uint64_t x = 0x9E3779B97F4A7C15ull ^ (uint64_t)round_;
        std::vector<int64_t> start(clauses_.max_len() + 2, 0);
        for (int len = 1; len <= clauses_.max_len(); len++)
            start[len + 1] = start[len] + clauses_.rows(len);
        for (long i = 0; i < want; i++) {
            x ^= x << 13;
            x ^= x >> 7;
            x ^= x << 17;
            const int64_t g = (int64_t)(x % (uint64_t)total);
            const int len = (int)(std::upper_bound(start.begin(), start.end(), g) - start.begin()) - 1;
            const uint32_t t = clauses_.tag(len, (uint32_t)(g - start[len]));
            if (t != kNoTag) tags.push_back(t);
        }
        std::sort(tags.begin(), tags.end());
        tags.erase(std::unique(tags.begin(), tags.end()), tags.end());
        if (!tags.empty())
            check_rc(gss_res_snapshot(res_, round_, (int64_t)tags.size(), tags.data()),
                     "gss_res_snapshot");
    } else if (phase == 0 && round_ >= 16) {
        float loss = 0.f;
        int64_t samples = 0;
        check_rc(gss_res_train(res_, round_, 5, 2, 2048, 0.02f, &loss, &samples), "gss_res_train");
    }
}

// Appends a committed batch to the device under fresh tags (device policies).
// When the device cannot take it (out of memory, or of tags) the batch stays
// the tail of its buckets in the mirror, is remembered, and the DB is reduced
// once the run in flight is delivered (reduce_db_on_device adds it then).
bool EngineClauseSharer::add_on_device(const ClauseMirror::Batch &b) {
    std::vector<int64_t> per(clauses_.max_len() + 1, 0);
    for (int32_t len : b.lens) per[len]++;
    bool tagged = (int64_t)free_tags_.size() >= (int64_t)b.lens.size();
    for (int len = 1; tagged && len <= clauses_.max_len(); len++)
        if (per[len]) clauses_.tag_tail(len, per[len], free_tags_);
    int rc = GSS_E_NOMEM;
    if (tagged) {
        std::vector<int64_t> seen(clauses_.max_len() + 1, 0);
        std::vector<uint32_t> tags(b.lens.size());
        for (size_t i = 0; i < b.lens.size(); i++) {
            const int len = b.lens[i];
            tags[i] = clauses_.tag(len, (uint32_t)(clauses_.rows(len) - per[len] + seen[len]++));
        }
        rc = gss_res_add(res_, (int64_t)b.lens.size(), b.lits.data(), b.lens.data(), tags.data(),
                         round_);
    }
    if (rc == GSS_OK) return true;
    if (rc != GSS_E_NOMEM) check_rc(rc, "gss_res_add");
    ran_out_of_memory_ = true;
    reduce_pending_ = true;
    failed_lits_.insert(failed_lits_.end(), b.lits.begin(), b.lits.end());
    failed_lens_.insert(failed_lens_.end(), b.lens.begin(), b.lens.end());
    return false;
}

void EngineClauseSharer::deliver(InFlight &run) {
    // the run's reports: (clause, solver word) ids from a plain sweep, or one
    // record per clause with its solver words from a step
    const gss_hit_id *ids = nullptr;
    const gss_hit_clause *cls = nullptr;
    int64_t n_ids = 0, n_cls = 0;
    if (run.plain) {
        ids = id_buf_.data();
        n_ids = (int64_t)id_buf_.size();
    } else if (run.records && clause_records_) {
        check_rc(gss_step_clauses_map(engine_, &cls, &n_cls), "gss_step_clauses_map");
    } else if (run.records) {
        check_rc(gss_step_ids_map(engine_, &ids, &n_ids), "gss_step_ids_map");
    }
    int64_t n = n_ids;
    for (int64_t i = 0; i < n_cls; i++) n += __builtin_popcountll(cls[i].words);
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
    global_stats_[gpuReports] += (unsigned long)n;
    // one activity bump per (clause, solver) report (GpuRunner.cu:375-378)
    for (int64_t i = 0; i < n_ids; i++) clauses_.bump(ids[i].len, ids[i].row);
    for (int64_t i = 0; i < n_cls; i++)
        for (int k = __builtin_popcountll(cls[i].words); k > 0; k--)
            clauses_.bump(cls[i].len, cls[i].row);
    Gauge g(global_stats_[timeSpentFillingReported], opts_.quickProf);
    const int S = (int)run.taken.size();
    std::vector<ClauseBatch *> batch(S, nullptr);
    for (int s = 0; s < S; s++) {
        if (run.taken[s].count <= 0) continue;
        batch[s] = reports_[s]->get_new();
        batch[s]->start_assig = run.taken[s].start_id;
        batch[s]->assig_count = run.taken[s].count;
    }
    auto add = [&](int s, int len, uint32_t row) {
        if (s >= S || !batch[s]) return;  // cannot happen: hits only in valid slots
        ClauseBatch &b = *batch[s];
        b.clauses.push_back({clauses_.meta(len, row).id, (int)b.lits.size()});
        const uint32_t *l = clauses_.row_lits(len, row);
        b.lits.insert(b.lits.end(), l, l + len);
    };
    for (int64_t i = 0; i < n_ids; i++) add(ids[i].word, ids[i].len, ids[i].row);
    for (int64_t i = 0; i < n_cls; i++)
        for (uint64_t w = cls[i].words; w; w &= w - 1)
            add(__builtin_ctzll(w), cls[i].len, cls[i].row);
    for (int s = 0; s < S; s++)
        if (batch[s]) reports_[s]->publish(batch[s]);
}

// One run in flight, as the reference's pipeline (GpuRunner.cu:211-261): the
// run in flight is waited for (its records start to travel to the host), the
// next run is staged and submitted, then the waited run's records are
// delivered while the new one is on the device.
void EngineClauseSharer::run_once(bool can_start) {
    InFlight prev;
    const bool had = in_flight_.active;
    if (had) {
        prev = std::move(in_flight_);
        in_flight_ = InFlight();
        collect(prev);
    }
    if (can_start) start_run(in_flight_);
    if (had) deliver(prev);
    if (reduce_pending_) {
        reduce_pending_ = false;
        reduce_db_now();
    }
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
    if (policy_ != Policy::Host) {
        reduce_db_on_device();
        global_stats_[gpuClauses] = (unsigned long)clauses_.clause_count();
        global_stats_[gpuClauseLengthSum] = (unsigned long)clauses_.length_sum();
        log(2, "c Done reducing gpu clause db, clause count is " +
                   std::to_string(clauses_.clause_count()));
        return;
    }
    // Reduce, then upload every bucket whose device rows differ from the
    // mirror's: the reduced ones first (they shrink and give memory back),
    // then those a failed append left behind (the mirror already holds that
    // batch, also in the lengths 1 and 2 the reduction never touches).  As
    // the reference (Clauses.cu reduceDb: reduce and retry until the copy
    // fits), an upload that still does not fit reduces again; when nothing
    // is left to reduce, the buckets the device cannot hold are emptied on
    // both sides, so device and mirror always agree.
    for (;;) {
        const std::vector<int> changed = clauses_.reduce(limits_.max_clause_size);
        std::vector<int> order(changed);
        for (int len : dirty_)
            if (std::find(changed.begin(), changed.end(), len) == changed.end()) order.push_back(len);
        dirty_.insert(changed.begin(), changed.end());
        int rc = GSS_OK;
        for (int len : order) {
            const auto &lits = clauses_.bucket_lits(len);
            rc = gss_db_set_bucket(engine_, len, (int64_t)(lits.size() / len), lits.data());
            if (rc == GSS_E_NOMEM) break;
            check_rc(rc, "gss_db_set_bucket");
            dirty_.erase(len);
        }
        if (rc == GSS_OK) break;
        ran_out_of_memory_ = true;
        log(1, "c GPU memory still short after a reduction: " + std::string(gss_last_error()));
        if (changed.empty()) {
            const std::vector<int> left(dirty_.begin(), dirty_.end());
            for (int len : left) {
                clauses_.drop_bucket(len);
                check_rc(gss_db_set_bucket(engine_, len, 0, nullptr), "gss_db_set_bucket");
                dirty_.erase(len);
            }
            break;
        }
    }
    global_stats_[gpuClauses] = (unsigned long)clauses_.clause_count();
    global_stats_[gpuClauseLengthSum] = (unsigned long)clauses_.length_sum();
    log(2, "c Done reducing gpu clause db, clause count is " +
               std::to_string(clauses_.clause_count()));
}

// Device policies: the device keeps about half of the clauses of length >= 3
// (activity median, or the learned score's), compacting in place; the mirror
// drops the same rows.  A batch the device could not take is then added again
// (its tags assigned now if they were missing); if it still does not fit the
// reduction repeats, and when nothing more goes the batch is dropped from the
// mirror, so device and mirror always agree.
void EngineClauseSharer::reduce_db_on_device() {
    const int pol = policy_ == Policy::Learned ? GSS_RES_POLICY_LEARNED : GSS_RES_POLICY_ACTIVITY;
    for (;;) {
        int64_t kept = 0;
        std::fill(evicted_.begin(), evicted_.end(), 0);
        check_rc(gss_res_reduce(res_, pol, round_, evicted_.data(), &kept), "gss_res_reduce");
        std::vector<uint32_t> freed;
        clauses_.compact_by_tags(evicted_.data(), freed);
        free_tags_.insert(free_tags_.end(), freed.begin(), freed.end());
        if (failed_lens_.empty()) return;
        ClauseMirror::Batch b;
        b.lits.swap(failed_lits_);
        b.lens.swap(failed_lens_);
        if (add_on_device(b)) return;
        if (!freed.empty()) continue;  // reduced something: try again
        // nothing left to reduce: the batch leaves the mirror
        std::vector<int64_t> per(clauses_.max_len() + 1, 0);
        for (int32_t len : failed_lens_) per[len]++;
        std::vector<uint32_t> dropped;
        for (int len = 1; len <= clauses_.max_len(); len++)
            if (per[len]) clauses_.drop_tail(len, per[len], dropped);
        free_tags_.insert(free_tags_.end(), dropped.begin(), dropped.end());
        failed_lits_.clear();
        failed_lens_.clear();
        log(1, "c GPU memory short with nothing left to reduce: a batch of clauses was dropped");
        return;
    }
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
        if (lits[i] < 0 || (unsigned)(lits[i] >> 1) >= GSS_MAX_VARS) return -1;
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

// One byte per variable of the sharer's count (setVarCount), as the caller
// sizes the buffer; a solver may have set variables beyond it.
void EngineClauseSharer::getCurrentAssignment(int solver, uint8_t *assig) {
    SolverWindow &w = *windows_[solver];
    const int n = var_count_;
    {
        std::lock_guard<std::mutex> g(w.lock);
        w.current_assignment(assig, n);
    }
    for (int l : to_unset_[solver])
        if ((l >> 1) < n) assig[l >> 1] = kValUndef;
}

GpuClauseSharer *makeGpuClauseSharerPtr(GpuClauseSharerOptions opts,
                                        std::function<void(const std::string &str)> logFunc) {
    return new EngineClauseSharer(opts, std::move(logFunc));
}

}  // namespace GpuShare
