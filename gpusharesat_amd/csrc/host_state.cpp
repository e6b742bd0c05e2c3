// Host-side state of the sharer that needs no GPU: solver assignment windows,
// the clause mirror with its activity policy, and the per-solver report
// queues (see sharer_impl.h for the reference files each part follows).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <stdexcept>

#include "sharer_impl.h"

namespace GpuShare {

namespace {
constexpr float kRescale = 1e19f;  // RESCALE_CONST, gpuShareLib/Clauses.cuh:39
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
        var_count_.store(n, std::memory_order_release);
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

void SolverWindow::current_assignment(uint8_t *out, int n) const {
    const int known = std::min<int>(n, (int)last_val_.size());
    if (known > 0) std::memcpy(out, last_val_.data(), (size_t)known);
    if (n > known) std::memset(out + known, kValUndef, (size_t)(n - known));
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

void ClauseMirror::drop_bucket(int len) {
    if (len < 1 || len >= (int)meta_.size()) return;
    const long n = (long)meta_[len].size();
    count_ -= n;
    len_sum_ -= n * len;
    meta_[len].clear();
    lits_[len].clear();
}

void ClauseMirror::compact_by_tags(const uint8_t *evicted, std::vector<uint32_t> &freed) {
    for (size_t len = 1; len < meta_.size(); len++) {
        std::vector<ClauseMeta> &meta = meta_[len];
        std::vector<uint32_t> &lits = lits_[len];
        size_t to = 0;
        for (size_t r = 0; r < meta.size(); r++) {
            const ClauseMeta &m = meta[r];
            if (m.tag != kNoTag && evicted[m.tag]) {
                freed.push_back(m.tag);
                continue;
            }
            if (to != r) {
                meta[to] = m;
                std::memmove(&lits[to * len], &lits[r * len], len * sizeof(uint32_t));
            }
            to++;
        }
        const long removed = (long)(meta.size() - to);
        count_ -= removed;
        len_sum_ -= removed * (long)len;
        meta.resize(to);
        lits.resize(to * len);
    }
}

bool ClauseMirror::tag_tail(int len, int64_t count, std::vector<uint32_t> &free) {
    if ((int64_t)free.size() < count || len >= (int)meta_.size() ||
        count > (int64_t)meta_[len].size())
        return false;
    std::vector<ClauseMeta> &meta = meta_[len];
    for (size_t r = meta.size() - (size_t)count; r < meta.size(); r++) {
        if (meta[r].tag == kNoTag) {
            meta[r].tag = free.back();
            free.pop_back();
        }
    }
    return true;
}

void ClauseMirror::drop_tail(int len, int64_t count, std::vector<uint32_t> &freed) {
    if (len >= (int)meta_.size()) return;
    std::vector<ClauseMeta> &meta = meta_[len];
    count = std::min<int64_t>(count, (int64_t)meta.size());
    for (size_t r = meta.size() - (size_t)count; r < meta.size(); r++)
        if (meta[r].tag != kNoTag) freed.push_back(meta[r].tag);
    meta.resize(meta.size() - (size_t)count);
    lits_[len].resize(meta.size() * len);
    count_ -= (long)count;
    len_sum_ -= (long)count * len;
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

}  // namespace GpuShare
