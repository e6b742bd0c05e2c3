// GpuShare::GpuClauseSharer implemented over the gss check engine.
//
// Host structures and the reference semantics they follow (GpuShareSat tree):
//   SolverWindow    one solver's assignment slots, set/unset/send, delta
//                   updates and roll-forward position
//                   (gpuShareLib/Assigs.cu:146-300, GpuClauseSharerImpl.cu:130-191)
//   ClauseMirror    host copy of the device clause buckets with per-clause id,
//                   lbd and activity; staging of new clauses; activity bump /
//                   decay / rescale; reduceDb by approximate activity median
//                   (gpuShareLib/Clauses.cu:200-291,318-356,391-525)
//   ReportQueues    per-solver batches of reported clauses, deduplication and
//                   own-clause suppression (gpuShareLib/Reported.cu:35-204)
//   run pipeline    one run in flight, as one batched step (gss_step_submit):
//                   its hits are collected by the next gpuRun()
//                   (gpuShareLib/GpuRunner.cu:211-383)
#pragma once

#include <atomic>
#include <cstdint>
#include <deque>
#include <memory>
#include <mutex>
#include <queue>
#include <set>
#include <string>
#include <vector>

#include "gpuShareLib/GpuClauseSharer.h"
#include "gss.h"
#include "gss_residency.h"

namespace GpuShare {

enum : uint8_t { kValTrue = 0, kValFalse = 1, kValUndef = 2 };

// Slot width and clause size limits.  Native: 64 slots per solver (one
// uint64 word), clauses up to 1024 literals.  Reference-compatible (env
// GPUSHARESAT_COMPAT=1): 32 slots and 100 literals, as the reference's
// Vals = uint32 and MAX_CL_SIZE (gpuShareLib/BaseTypes.cuh:28-32).
struct SharerLimits {
    int slots = 64;
    int max_clause_size = 1024;
    static SharerLimits from_env();
};

class SolverWindow {
public:
    std::mutex lock;

    explicit SolverWindow(int slots);
    void set_var_count(int n);
    void set_var(int var, uint8_t val);
    bool available() const { return current_id_ != first_id_ + slots_; }
    long done();
    // values of variables [0, n): those the window never saw are undefined
    void current_assignment(uint8_t *out, int n) const;
    // readable without the lock (the GPU thread sizes the table from it)
    int var_count() const { return var_count_.load(std::memory_order_acquire); }

    struct Taken {
        long start_id = 0;
        long count = 0;
        uint64_t valid = 0;
        int last_pos = -1;  // -1: nothing to roll forward
    };
    // Moves this window's updates (as word `word`) to out and opens the next
    // window (copyUpdatesLocked, Assigs.cu:243-300).
    Taken take(std::vector<gss_var_update> &out, int word);

private:
    int slots_;
    uint64_t slot_bits_;
    uint64_t not_completed_;
    long current_id_ = 0;
    long first_id_ = 0;
    std::vector<uint8_t> last_val_;
    std::atomic<int> var_count_{0};  // last_val_.size(), published after a resize
    std::vector<int> update_pos_;
    std::vector<gss_var_update> updates_;
};

constexpr uint32_t kNoTag = 0xFFFFFFFFu;

struct ClauseMeta {
    long id;
    int lbd;
    float activity;
    uint32_t tag = kNoTag;  // slot in the device residency state (device policies)
};

class ClauseMirror {
public:
    explicit ClauseMirror(float decay) : decay_(decay) {}
    void set_decay(float decay) { decay_ = decay; }
    // any thread
    long stage(const int *lits, int count, int lbd);
    // GPU thread: moves staged clauses into the buckets; returns them grouped
    // as (lits, lens) for the device append and their rows.
    struct Batch {
        std::vector<uint32_t> lits;
        std::vector<int32_t> lens;
        std::vector<uint32_t> rows;
        int64_t max_lit = -1;
    };
    void commit_staged(Batch &b);
    void bump(int len, uint32_t row);
    // keeps clauses per the reference policy; returns the buckets changed
    std::vector<int> reduce(int max_lbd);
    // removes every clause of bucket len
    void drop_bucket(int len);
    // Device policies: keeps the rows whose tag is not flagged in evicted
    // (rows without a tag are kept), in order, as the device compaction
    // does; the tags of the removed rows go to freed.
    void compact_by_tags(const uint8_t *evicted, std::vector<uint32_t> &freed);
    // Tags for the last count rows of bucket len (in order) from free.
    bool tag_tail(int len, int64_t count, std::vector<uint32_t> &free);
    // Removes the last count rows of bucket len; their tags go to freed.
    void drop_tail(int len, int64_t count, std::vector<uint32_t> &freed);
    uint32_t tag(int len, uint32_t row) const { return meta_[len][row].tag; }
    float approx_nth_activity(int min_lbd, int max_lbd, long n) const;

    long clause_count() const { return count_; }
    long length_sum() const { return len_sum_; }
    long added() const { return added_.load(); }
    int max_len() const { return (int)lits_.size() - 1; }
    int64_t rows(int len) const {
        return len < (int)meta_.size() ? (int64_t)meta_[len].size() : 0;
    }
    const uint32_t *row_lits(int len, uint32_t row) const { return &lits_[len][(size_t)row * len]; }
    const ClauseMeta &meta(int len, uint32_t row) const { return meta_[len][row]; }
    const std::vector<uint32_t> &bucket_lits(int len) const { return lits_[len]; }
    bool has_staged();
    float act_incr() const { return act_incr_; }
    bool find(long id, float &activity, int &lbd) const {
        for (const auto &bucket : meta_)
            for (const ClauseMeta &m : bucket)
                if (m.id == id) {
                    activity = m.activity;
                    lbd = m.lbd;
                    return true;
                }
        return false;
    }

private:
    void rescale();
    std::mutex stage_lock_;
    std::vector<int32_t> staged_lits_;
    std::vector<int> staged_lens_;
    std::vector<int> staged_lbd_;
    std::vector<long> staged_ids_;
    long next_id_ = 0;
    std::atomic<long> added_{0};

    std::vector<std::vector<uint32_t>> lits_;  // [len] row-major
    std::vector<std::vector<ClauseMeta>> meta_;
    long count_ = 0, len_sum_ = 0;
    float act_incr_ = 1.0f;
    float decay_;
};

struct ClauseBatch {
    std::vector<int> lits;
    std::vector<std::pair<long, int>> clauses;  // (gpuClauseId, start in lits)
    size_t next = 0;
    long start_assig = 0, assig_count = 0;
    long knows_from = 0;  // first assignment id that knew about these clauses
    void clear() {
        lits.clear();
        clauses.clear();
        next = 0;
        start_assig = assig_count = 0;
        knows_from = 0;
    }
};

class ReportQueue {
public:
    // producer (GPU thread)
    ClauseBatch *get_new();
    void publish(ClauseBatch *b);
    // consumer (the solver's thread)
    bool pop(int *&lits, int &count, long &id, std::atomic<unsigned long> *stats);
    void clause_was_added(long id);
    void assig_was_sent(long id) { last_sent_ = id; }
    long last_all_reported() const { return last_all_reported_.load(); }
    ~ReportQueue();

private:
    std::mutex m_;
    std::deque<ClauseBatch *> ready_;
    std::vector<ClauseBatch *> pool_;
    // consumer side
    ClauseBatch *current_ = nullptr;
    std::deque<ClauseBatch *> taken_;
    std::set<long> no_import_;
    std::queue<std::pair<long, long>> dont_import_;  // (clause id, assig id)
    std::atomic<long> last_sent_{0};
    std::atomic<long> last_all_reported_{0};
};

class EngineClauseSharer : public GpuClauseSharer {
public:
    EngineClauseSharer(GpuClauseSharerOptions opts,
                       std::function<void(const std::string &)> log_func);
    ~EngineClauseSharer() override;

    void gpuRun() override;
    void reduceDb() override;
    long getAddedClauseCount() override;
    long getAddedClauseCountAtLastReduceDb() override;
    bool hasRunOutOfGpuMemoryOnce() override;
    void getGpuMemInfo(size_t &free, size_t &total) override;
    int getGlobalStatCount() override;
    long getGlobalStat(GlobalStats stat) override;
    void writeClausesInCnf(FILE *file) override;
    void setVarCount(int newCount) override;
    long addClause(int solverId, int *lits, int count) override;
    void setCpuSolverCount(int count) override;
    const char *getOneSolverStatName(OneSolverStats stat) override;
    const char *getGlobalStatName(GlobalStats stat) override;
    int getOneSolverStatCount() override;
    bool trySetSolverValues(int cpuSolverId, int *lits, int count) override;
    void unsetSolverValues(int cpuSolverId, int *lits, int count) override;
    long trySendAssignment(int cpuSolverId) override;
    bool popReportedClause(int cpuSolverId, int *&lits, int &count, long &gpuClauseId) override;
    long getLastAssigAllReported(int cpuSolverId) override;
    void getCurrentAssignment(int cpuSolverId, uint8_t *assig) override;
    long getOneSolverStat(int cpuSolverId, OneSolverStats stat) override;

    // extras for the C ABI / tests
    int slots_per_solver() const { return limits_.slots; }
    int max_clause_size() const { return limits_.max_clause_size; }
    gss_engine *engine() { return engine_; }
    int solver_count() const { return (int)windows_.size(); }
    bool debug_clause(long id, float &activity, int &lbd) const {
        return clauses_.find(id, activity, lbd);
    }

private:
    struct InFlight {
        bool active = false;
        std::vector<SolverWindow::Taken> taken;  // per solver
        std::vector<gss_var_update> updates;
        std::vector<int8_t> last_pos;
        long clause_count = 0;
        int64_t records = 0;  // hit records of the waited step
        bool plain = false;   // a plain sweep (device policies) instead of a step
        bool swept = false;   // plain: a sweep was launched (some slot was valid)
        bool idle = false;    // nothing to append, apply or check: no device work
        bool screened = false;  // the step was submitted with GSS_STEP_SCREEN
    };

    void run_once(bool can_start);
    void reduce_db_now();
    void reduce_db_on_device();
    bool add_on_device(const ClauseMirror::Batch &b);
    void train_policy();
    bool start_run(InFlight &run);
    void collect(InFlight &run);
    void deliver(InFlight &run);
    void unset_pending_locked(int solver);
    void ensure_table();
    void log(int level, const std::string &msg);
    int words_for(int solvers) const;

    GpuClauseSharerOptions opts_;
    std::function<void(const std::string &)> log_func_;
    SharerLimits limits_;
    gss_engine *engine_ = nullptr;
    int var_count_ = 0;
    int64_t table_vars_ = 0;
    int table_words_ = 0;
    int64_t clause_vars_ = 0;  // 1 + largest variable of any clause on the device

    ClauseMirror clauses_;
    std::vector<std::unique_ptr<SolverWindow>> windows_;
    std::vector<std::unique_ptr<ReportQueue>> reports_;
    std::vector<std::vector<int>> to_unset_;

    std::vector<std::atomic<unsigned long>> global_stats_;
    std::vector<std::unique_ptr<std::atomic<unsigned long>[]>> solver_stats_;
    std::vector<std::string> global_names_, solver_names_;

    InFlight in_flight_;
    long added_at_last_reduce_ = 0;
    bool ran_out_of_memory_ = false;
    bool reduce_pending_ = false;  // out of memory: reduce once the run in flight is delivered
    std::set<int> dirty_;          // buckets whose device rows differ from the mirror's

    // Reduction policy (GPUSHARESAT_REDUCE_POLICY): the reference's activity
    // median on the host mirror (default), or on the device over a tagged DB,
    // or the learned score (gss_residency.h).  The device policies run each
    // gpuRun as a plain sweep whose records feed the per-clause state.
    enum class Policy { Host, Device, Learned };
    Policy policy_ = Policy::Host;
    bool clause_records_ = true;  // GPUSHARESAT_RECORDS: clauses (default) or ids
    // Runs as GSS_STEP_SCREEN steps (same reports): GPUSHARESAT_SCREEN 1 =
    // always, 0 = never, unset = chosen from the runs' device times.  In
    // that mode the last kScreenProbe of every screen_period_ runs use the
    // other form; each form averages the device ms per resident clause of
    // its runs in the period that followed a run of the same form (a switch
    // makes the next screened step rebuild the aggregate), and at the end of
    // a period the other form takes over if it was at least 5 % faster.
    // Screened first.
    enum class Screen { Off, On, Auto };
    Screen screen_cfg_ = Screen::Auto;
    static constexpr int kScreenProbe = 4;
    int screen_period_ = 256;  // GPUSHARESAT_SCREEN=auto:N sets it (N >= 8)
    int64_t screen_runs_ = 0;
    bool screen_best_ = true;
    double screen_sum_[2] = {0.0, 0.0};  // [unscreened, screened] over the period
    int64_t screen_n_[2] = {0, 0};
    int last_screened_ = -1;             // form of the last step run (-1: none yet)
    bool choose_screen();
    void note_screen_time(bool screened, float ms, long clauses);
    gss_res *res_ = nullptr;
    int64_t max_tags_ = 0;
    std::vector<uint32_t> free_tags_;
    std::vector<uint8_t> evicted_;
    int32_t round_ = 0;
    std::vector<gss_hit> hit_buf_;
    std::vector<gss_hit_id> id_buf_;
    // a batch the device could not take (no memory, or no free tags): it is
    // the tail of its buckets in the mirror, its clause order is kept here
    std::vector<uint32_t> failed_lits_;
    std::vector<int32_t> failed_lens_;
};

}  // namespace GpuShare
