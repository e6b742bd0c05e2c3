// The serving loop of include/gss_serve.h: request and patch rings in
// fine-grained pinned host memory, a serving thread that keeps batched steps
// on the device (each admitting published requests on the device when it
// starts), and per-request results.  Extends the reference's run loop
// (GpuClauseSharerImpl.cu:83-94, GpuRunner.cu:211-261) from "the assignments
// sent since the last run" to requests arriving at any time.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "engine_internal.h"
#include "errors.h"
#include "gss_serve.h"

using gss::kServeMaxRequests;
using gss::ServeReport;
using gss::ServeReq;

namespace {

using Clock = std::chrono::steady_clock;

// Pinned record buffers per engine step slot: a step's records are lent from
// them for 2 * kRecordBuffers steps (two slots), so a client that releases a
// result within a few steps never costs a copy.
constexpr int kRecordBuffers = 4;

int fail(int code, const std::string &msg) { return gss::record_error(code, msg.c_str()); }

// A published request, until its step completes.
struct Pending {
    uint64_t seq;
    int solver;
    int k;
    uint64_t patch_off;
    uint32_t n_patches;
    Clock::time_point t_submit;
};

// A launched step, until it completes.
struct StepInfo {
    int64_t index = 0;
    int64_t clauses_after = 0;   // clauses in the DB once the step appended its batch
    int64_t literals_after = 0;
    int64_t bups_through = 0;    // base updates staged up to and including this step
    uint64_t published_at_launch = 0;
};

// One step's records grouped by request, as the step recorded them (rw
// 8-byte words each; the first word is row | len << 32 | word << 48 with the
// word made relative to the request).
//
// Grouped on the device, a step's records are lent straight from the engine's
// pinned buffers (valid for 2 * kRecordBuffers steps); before a step that may
// reuse the buffer is waited for, the records of a step whose results a
// client still holds unread are copied (materialize).
struct StepRecords {
    std::vector<uint64_t> words;
    const uint64_t *lent = nullptr;  // engine pinned memory, while valid
    int64_t n = 0;                   // records
    int rw = 3;
    bool absolute = false;           // words not yet made relative (subtract first_word)
    const uint64_t *data() const { return lent ? lent : words.data(); }
    void materialize() {
        if (!lent) return;
        words.assign(lent, lent + (size_t)n * rw);
        lent = nullptr;
    }
};

struct Done {
    gss_serve_result r{};
    std::shared_ptr<const StepRecords> step;
    size_t first = 0;                 // its first record in step->words
    std::vector<gss_serve_hit> recs;  // translated on request (gss_server_records)
    bool translated = false;
};

}  // namespace

struct gss_server {
    gss_engine *e = nullptr;
    gss_server_config cfg{};
    ServeReq *reqs = nullptr;        // pinned, fine-grained
    gss_var_patch *patches = nullptr;
    uint64_t *head = nullptr;        // requests published (read by the device)

    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    uint64_t published = 0, patch_published = 0;
    uint64_t completed = 0, patch_completed = 0;  // requests / patches whose step completed
    std::deque<Pending> pending;                  // published, not completed, in order
    int64_t pending_words = 0;                    // table words those requests take
    std::vector<gss_base_update> bups;            // not yet staged
    std::vector<uint64_t> seen;                   // [base][var / 64] bits, for deduplication
    int64_t bups_submitted = 0;
    std::vector<uint32_t> cl_lits;                // admitted, not yet staged
    std::vector<int32_t> cl_lens;
    int64_t clauses_admitted = 0, literals_admitted = 0, clauses_staged = 0, literals_staged = 0;
    int64_t refused = 0;
    std::vector<std::vector<int64_t>> ids;        // [length][row] -> clause id
    std::deque<StepInfo> inflight;
    int64_t steps = 0;
    std::unordered_map<int64_t, Done> done;
    std::deque<std::pair<int64_t, std::weak_ptr<StepRecords>>> lent;  // (step, records)
    std::deque<int64_t> ready;
    int64_t records = 0, requests_done = 0;
    double checks = 0.0, device_ms = 0.0;
    int64_t host_grouped = 0;  // steps grouped by this thread (under mu)
    // host time of the serving thread's step submissions and waits (under mu)
    double submit_ms_total = 0.0, submit_ms_max = 0.0, wait_ms_max = 0.0;
    bool stop = false;
    int err_code = 0;
    std::string err;
    std::thread th;
};

namespace {

bool can_launch(const gss_server *s) {
    if (s->err_code || s->stop || (int)s->inflight.size() >= s->cfg.max_inflight) return false;
    // requests published that no launched step can have seen yet, or more
    // words of requests than the steps in flight can admit
    return s->published > s->completed &&
           (s->inflight.empty() || s->published > s->inflight.back().published_at_launch ||
            s->pending_words > (int64_t)s->inflight.size() * s->cfg.words);
}

// can_launch once the oldest step (which admitted `done` requests) is off
// the device: it no longer counts as in flight.
bool can_launch_after(const gss_server *s, int done) {
    if (s->err_code || s->stop || (int)s->inflight.size() - 1 >= s->cfg.max_inflight) return false;
    const uint64_t completed = s->completed + (uint64_t)done;
    if (s->published <= completed) return false;
    if (s->inflight.size() == 1) return true;  // only the finished step was in flight
    int64_t words_done = 0;
    for (int i = 0; i < done; i++) words_done += (s->pending[i].k + 63) / 64;
    return s->published > s->inflight.back().published_at_launch ||
           s->pending_words - words_done > (int64_t)(s->inflight.size() - 1) * s->cfg.words;
}

void set_failed(gss_server *s, int code, const char *what) {
    s->err_code = code ? code : GSS_E_STATE;
    s->err = std::string(what) + ": " + gss_last_error();
    s->cv_done.notify_all();
}

// Stages the pending base updates and clauses and submits a serving step.
void launch(gss_server *s, std::unique_lock<std::mutex> &lk) {
    StepInfo si;
    si.index = s->steps++;
    // one update per (base, variable) in a step: the last one submitted wins
    std::vector<gss_base_update> bu;
    bu.swap(s->bups);
    if (bu.size() > 1) {
        // newest first: a (base, variable) already seen has a later update
        const size_t vw = (size_t)(s->cfg.var_count + 63) / 64;
        if (s->seen.empty()) s->seen.assign((size_t)s->cfg.solvers * vw, 0ull);
        std::vector<gss_base_update> kept;
        kept.reserve(bu.size());
        for (size_t j = bu.size(); j-- > 0;) {
            uint64_t &w = s->seen[(size_t)bu[j].base * vw + bu[j].var / 64];
            const uint64_t bit = 1ull << (bu[j].var % 64);
            if (w & bit) continue;
            w |= bit;
            kept.push_back(bu[j]);
        }
        for (const gss_base_update &u : kept) s->seen[(size_t)u.base * vw + u.var / 64] = 0ull;
        bu.swap(kept);
    }
    si.bups_through = s->bups_submitted;
    std::vector<uint32_t> lits;
    std::vector<int32_t> lens;
    lits.swap(s->cl_lits);
    lens.swap(s->cl_lens);
    const int64_t first_clause = s->clauses_staged;
    s->clauses_staged += (int64_t)lens.size();
    s->literals_staged += (int64_t)lits.size();
    si.clauses_after = s->clauses_staged;
    si.literals_after = s->literals_staged;
    si.published_at_launch = s->published;
    lk.unlock();
    std::vector<uint32_t> rows(lens.size());
    uint64_t valid[GSS_MAX_WORDS] = {};
    gss_step_desc d{};
    d.n_clauses = (int64_t)lens.size();
    d.lits = lits.data();
    d.lens = lens.data();
    d.out_rows = rows.data();
    d.n_base_ups = (int64_t)bu.size();
    d.base_ups = bu.data();
    d.valid = valid;
    d.records = s->cfg.records;
    const Clock::time_point t_sub = Clock::now();
    const int rc = gss::step_submit_ex(s->e, &d, true);
    const double sub_ms = std::chrono::duration<double, std::milli>(Clock::now() - t_sub).count();
    lk.lock();
    s->submit_ms_total += sub_ms;
    s->submit_ms_max = std::max(s->submit_ms_max, sub_ms);
    gss::log_stall("serving step submit", sub_ms);
    if (rc) {
        set_failed(s, rc, "serving step");
        return;
    }
    for (size_t i = 0; i < lens.size(); i++) {
        const int len = lens[i];
        if ((int)s->ids.size() <= len) s->ids.resize(len + 1);
        std::vector<int64_t> &v = s->ids[len];
        if (v.size() <= rows[i]) v.resize((size_t)rows[i] + 1, -1);
        v[rows[i]] = first_clause + (int64_t)i;
    }
    s->inflight.push_back(si);
}

// Waits for the oldest step and hands each request it admitted its records:
// two sequential passes over the step's records in pinned memory (count per
// request, then copy grouped by request), outside the lock; clause ids are
// looked up only when a client reads a request's records.
void complete_oldest(gss_server *s, std::unique_lock<std::mutex> &lk) {
    // the step about to be waited for may reuse the pinned record buffer of
    // the step 2 * kRecordBuffers before it (its slot's kRecordBuffers-th
    // last wait with records, at the earliest): records lent from there that a
    // client still holds are copied now
    const int64_t next = s->inflight.front().index;
    while (!s->lent.empty() && s->lent.front().first <= next - 2 * kRecordBuffers) {
        if (auto sp = s->lent.front().second.lock()) sp->materialize();
        s->lent.pop_front();
    }
    lk.unlock();
    int64_t n = 0;
    float ms = 0.f;
    const Clock::time_point t_wait = Clock::now();
    int rc = gss_step_wait(s->e, &n, &ms);
    const double wait_ms =
        std::chrono::duration<double, std::milli>(Clock::now() - t_wait).count();
    // what the step admitted and where its records of each word start (both
    // written by the step's graph into pinned memory) are read before the
    // next step is launched: it reuses this slot's pinned header
    ServeReport rep{};
    if (!rc) rc = gss::step_serve_report(s->e, &rep);
    uint64_t goff[GSS_MAX_WORDS + 1];
    bool grouped = false;
    if (!rc && n > 0) rc = gss::step_word_groups(s->e, goff, &grouped);
    const Clock::time_point now = Clock::now();
    // the requests it admitted are the oldest pending ones (only this thread
    // removes them, so they stay put while the lock is released)
    const int cnt = rc ? 0 : (int)rep.count;
    std::vector<int> ks(cnt);
    bool order_ok;
    {
        lk.lock();
        order_ok = rep.first == s->completed && (size_t)cnt <= s->pending.size();
        for (int i = 0; order_ok && i < cnt; i++) ks[i] = s->pending[i].k;
        // keep the device busy: the next step goes out before this one's
        // records are grouped (they stay in the slot's pinned memory until
        // that slot is waited for again, two steps on)
        if (!rc && order_ok && (int)s->inflight.size() <= s->cfg.max_inflight &&
            can_launch_after(s, cnt))
            launch(s, lk);
        lk.unlock();
    }
    // the records themselves: waits for their copy to pinned memory, which
    // the step just launched does not touch (its graph waits for that copy
    // on the device before it reuses the slot's device buffers)
    const uint64_t *raw = nullptr;
    const int rw = s->cfg.records == GSS_RECORDS_IDS ? 1 : 3;
    if (!rc && n > 0) {
        if (rw == 1) {
            const gss_hit_id *p = nullptr;
            rc = gss_step_ids_map(s->e, &p, &n);
            raw = (const uint64_t *)p;
        } else {
            const gss_hit *p = nullptr;
            rc = gss_step_hits_map(s->e, &p, &n);
            raw = (const uint64_t *)p;
        }
    }
    int owner[GSS_MAX_WORDS];
    uint32_t w0[kServeMaxRequests];
    std::fill(owner, owner + GSS_MAX_WORDS, -1);
    uint32_t w = 0;
    double assigs = 0;
    for (int i = 0; order_ok && i < cnt; i++) {
        const uint32_t nw = (uint32_t)(ks[i] + 63) / 64;
        w0[i] = w;
        for (uint32_t j = 0; j < nw && w + j < (uint32_t)GSS_MAX_WORDS; j++) owner[w + j] = i;
        w += nw;
        assigs += ks[i];
    }
    auto step = std::make_shared<StepRecords>();
    step->rw = rw;
    std::vector<size_t> first(cnt + 1, 0);
    std::vector<size_t> last(cnt, 0);
    bool grouped_here = false;  // the records outgrew the device grouping
    if (order_ok && n > 0 && grouped) {
        // grouped by word on the device: each request's records are one
        // range of the engine's pinned buffer, lent until the slot's reuse
        step->lent = raw;
        step->n = n;
        step->absolute = true;
        for (int i = 0; i < cnt; i++) {
            const uint32_t nw = (uint32_t)(ks[i] + 63) / 64;
            first[i] = goff[w0[i]];
            last[i] = goff[std::min<uint32_t>(w0[i] + nw, GSS_MAX_WORDS)];
        }
    } else if (order_ok && n > 0) {
        grouped_here = true;
        std::vector<size_t> count(cnt, 0);
        for (int64_t r = 0; r < n; r++) {
            const uint32_t word = (uint32_t)(raw[r * rw] >> 48);
            const int i = word < (uint32_t)GSS_MAX_WORDS ? owner[word] : -1;
            if (i >= 0) count[i]++;  // other words are never valid
        }
        for (int i = 0; i < cnt; i++) first[i + 1] = first[i] + count[i];
        for (int i = 0; i < cnt; i++) last[i] = first[i + 1];
        step->words.resize(first[cnt] * rw);
        std::vector<size_t> at(first.begin(), first.end() - 1);
        for (int64_t r = 0; r < n; r++) {
            const uint64_t id = raw[r * rw];
            const uint32_t word = (uint32_t)(id >> 48);
            const int i = word < (uint32_t)GSS_MAX_WORDS ? owner[word] : -1;
            if (i < 0) continue;
            uint64_t *dst = step->words.data() + at[i]++ * rw;
            dst[0] = (id & 0x0000FFFFFFFFFFFFull) | ((uint64_t)(word - w0[i]) << 48);
            for (int q = 1; q < rw; q++) dst[q] = raw[r * rw + q];
        }
    } else {
        std::fill(first.begin(), first.end(), 0);
    }
    lk.lock();
    if (grouped_here) s->host_grouped++;
    s->wait_ms_max = std::max(s->wait_ms_max, wait_ms);
    gss::log_stall("serving step wait", wait_ms);
    const StepInfo si = s->inflight.front();
    s->inflight.pop_front();
    if (rc) {
        set_failed(s, rc, "serving step wait");
        return;
    }
    if (!order_ok) {
        s->err_code = GSS_E_STATE;
        s->err = "serving step admitted requests out of order";
        s->cv_done.notify_all();
        return;
    }
    for (int i = 0; i < cnt; i++) {
        const Pending &p = s->pending[i];
        Done &dn = s->done[(int64_t)p.seq];
        dn.r.request = (int64_t)p.seq;
        dn.r.step = si.index;
        dn.r.clauses = si.clauses_after;
        dn.r.base_updates = si.bups_through;
        dn.r.solver = p.solver;
        dn.r.k = p.k;
        dn.r.first_word = (int32_t)w0[i];
        dn.r.latency_ms = std::chrono::duration<double, std::milli>(now - p.t_submit).count();
        dn.r.n_records = (int64_t)(last[i] - first[i]);
        dn.r.records = nullptr;
        dn.step = step;
        dn.first = first[i];
        s->pending_words -= (p.k + 63) / 64;
        s->ready.push_back(dn.r.request);
    }
    if (step->lent) s->lent.emplace_back(si.index, step);
    if (cnt) {
        const Pending &last = s->pending[cnt - 1];
        s->patch_completed = last.patch_off + last.n_patches;
    }
    s->pending.erase(s->pending.begin(), s->pending.begin() + cnt);
    s->completed += (uint64_t)cnt;
    s->records += n;
    s->requests_done += cnt;
    s->device_ms += ms;
    s->checks += (double)si.literals_after * assigs;
    s->cv_done.notify_all();
}

void serve_loop(gss_server *s) {
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        s->cv_work.wait(lk, [&] {
            return s->stop || s->err_code || can_launch(s) || !s->inflight.empty();
        });
        if (s->err_code) break;
        if (can_launch(s)) {
            launch(s, lk);
            continue;
        }
        if (!s->inflight.empty()) {
            complete_oldest(s, lk);
            continue;
        }
        if (s->stop) break;
    }
    // a failed server still drains its steps, so the device is idle at exit;
    // the records lent from the engine are copied before any slot is reused
    for (auto &l : s->lent)
        if (auto sp = l.second.lock()) sp->materialize();
    s->lent.clear();
    while (!s->inflight.empty()) {
        lk.unlock();
        int64_t n;
        float ms;
        (void)gss_step_wait(s->e, &n, &ms);
        lk.lock();
        s->inflight.pop_front();
    }
}

int check(gss_server *s) { return s ? GSS_OK : fail(GSS_E_INVAL, "null server"); }

void destroy(gss_server *s) {
    if (s->th.joinable()) {
        {
            std::lock_guard<std::mutex> g(s->mu);
            s->stop = true;
        }
        s->cv_work.notify_all();
        s->th.join();
    }
    if (s->e) gss_engine_destroy(s->e);  // drains the engine stream first
    if (s->reqs) (void)hipHostFree(s->reqs);
    if (s->patches) (void)hipHostFree(s->patches);
    if (s->head) (void)hipHostFree(s->head);
    delete s;
}

}  // namespace

extern "C" {

int gss_server_default_config(gss_server_config *c) {
    if (!c) return fail(GSS_E_INVAL, "null config");
    std::memset(c, 0, sizeof *c);
    c->var_count = 0;
    c->solvers = 1;
    c->words = 64;
    c->hbm_budget = 120e9;
    c->max_inflight = 2;
    c->records = GSS_RECORDS_FULL;
    c->request_ring = 4096;
    c->patch_ring = 8 << 20;
    return GSS_OK;
}

int gss_server_create(int device, const gss_server_config *c, gss_server **out) {
    if (!c || !out) return fail(GSS_E_INVAL, "null argument");
    *out = nullptr;
    if (c->var_count < 1 || c->solvers < 1 || c->solvers > 65535 || c->words < 1 ||
        c->words > GSS_MAX_WORDS || (c->words & (c->words - 1)) || c->max_inflight < 1 ||
        c->max_inflight > 2 || (c->records != GSS_RECORDS_FULL && c->records != GSS_RECORDS_IDS) ||
        c->request_ring < 1 || c->patch_ring < 1 || !(c->hbm_budget > 0))
        return fail(GSS_E_INVAL, "bad server configuration");
    gss_server *s = new (std::nothrow) gss_server();
    if (!s) return fail(GSS_E_NOMEM, "server allocation failed");
    s->cfg = *c;
    int rc;
    if ((rc = gss_engine_create(device, &s->e)) || (rc = gss_assig_reset(s->e, c->var_count, c->words)) ||
        (rc = gss_bases_reset(s->e, c->solvers))) {
        const std::string m = gss_last_error();
        destroy(s);
        return fail(rc, m);
    }
    // rings the host writes while the device may read them: fine-grained
    const unsigned fl = hipHostMallocCoherent;
    if (hipHostMalloc((void **)&s->reqs, (size_t)c->request_ring * sizeof(ServeReq), fl) !=
            hipSuccess ||
        hipHostMalloc((void **)&s->patches, (size_t)c->patch_ring * sizeof(gss_var_patch), fl) !=
            hipSuccess ||
        hipHostMalloc((void **)&s->head, 64, fl) != hipSuccess) {
        destroy(s);
        return fail(GSS_E_NOMEM, "pinned ring allocation failed");
    }
    __atomic_store_n(s->head, (uint64_t)0, __ATOMIC_RELEASE);
    // records lent to clients stay in the engine's pinned buffers for
    // 2 * kRecordBuffers steps before a held one has to be copied
    if ((rc = gss::engine_set_record_buffers(s->e, kRecordBuffers)) ||
        (rc = gss::engine_serve_attach(s->e, s->reqs, (uint32_t)c->request_ring, s->patches,
                                       (uint64_t)c->patch_ring, s->head))) {
        const std::string m = gss_last_error();
        destroy(s);
        return fail(rc, m);
    }
    s->th = std::thread(serve_loop, s);
    *out = s;
    return GSS_OK;
}

int gss_server_destroy(gss_server *s) {
    if (!s) return GSS_OK;
    destroy(s);
    return GSS_OK;
}

int gss_server_engine(gss_server *s, gss_engine **out) {
    int rc = check(s);
    if (rc) return rc;
    if (!out) return fail(GSS_E_INVAL, "null out");
    *out = s->e;
    return GSS_OK;
}

int gss_server_update_base(gss_server *s, int64_t n, const gss_base_update *ups) {
    int rc = check(s);
    if (rc) return rc;
    if (n < 0 || (n && !ups)) return fail(GSS_E_INVAL, "bad base updates");
    for (int64_t i = 0; i < n; i++)
        if (ups[i].var >= (uint64_t)s->cfg.var_count || ups[i].base >= s->cfg.solvers)
            return fail(GSS_E_INVAL, "base update " + std::to_string(i) + " out of range");
    std::lock_guard<std::mutex> g(s->mu);
    if (s->err_code) return fail(s->err_code, s->err);
    s->bups.insert(s->bups.end(), ups, ups + n);
    s->bups_submitted += n;
    return GSS_OK;
}

int gss_server_add_clauses(gss_server *s, int64_t n, const uint32_t *lits, const int32_t *lens,
                           int64_t *first_id) {
    int rc = check(s);
    if (rc) return rc;
    if (n < 0 || (n && (!lits || !lens)) || !first_id) return fail(GSS_E_INVAL, "bad clauses");
    *first_id = -1;
    int64_t total = 0;
    for (int64_t i = 0; i < n; i++) {
        if (lens[i] < 1 || lens[i] > GSS_MAX_CLAUSE_LEN)
            return fail(GSS_E_INVAL, "clause " + std::to_string(i) + ": bad length");
        total += lens[i];
    }
    for (int64_t i = 0; i < total; i++)
        if ((int64_t)(lits[i] >> 1) >= s->cfg.var_count)
            return fail(GSS_E_INVAL, "literal " + std::to_string(lits[i]) + " past the variables");
    std::lock_guard<std::mutex> g(s->mu);
    if (s->err_code) return fail(s->err_code, s->err);
    if (4.0 * (double)(s->literals_admitted + total) > s->cfg.hbm_budget) {
        s->refused += n;
        return GSS_OK;  // over the budget: refused, *first_id = -1
    }
    *first_id = s->clauses_admitted;
    s->cl_lits.insert(s->cl_lits.end(), lits, lits + total);
    s->cl_lens.insert(s->cl_lens.end(), lens, lens + n);
    s->clauses_admitted += n;
    s->literals_admitted += total;
    return GSS_OK;
}

int gss_server_submit(gss_server *s, int solver, int k, int64_t n_patches,
                      const gss_var_patch *patches, int64_t *out_id) {
    int rc = check(s);
    if (rc) return rc;
    if (!out_id || n_patches < 0 || (n_patches && !patches))
        return fail(GSS_E_INVAL, "bad request");
    *out_id = -1;
    if (solver < 0 || solver >= s->cfg.solvers || k < 1 || k > 64 * s->cfg.words)
        return fail(GSS_E_INVAL, "request solver or size out of range");
    const uint32_t nw = (uint32_t)(k + 63) / 64;
    for (int64_t i = 0; i < n_patches; i++)
        if (patches[i].var >= (uint64_t)s->cfg.var_count || patches[i].word >= nw)
            return fail(GSS_E_INVAL, "patch " + std::to_string(i) + " out of range");
    {
        std::lock_guard<std::mutex> g(s->mu);
        if (s->err_code) return fail(s->err_code, s->err);
        if (s->stop) return fail(GSS_E_STATE, "server stopping");
        if (s->published - s->completed >= (uint64_t)s->cfg.request_ring)
            return fail(GSS_E_STATE, "request ring full");
        if (s->patch_published + (uint64_t)n_patches - s->patch_completed >
            (uint64_t)s->cfg.patch_ring)
            return fail(GSS_E_STATE, "patch ring full");
        const uint64_t pcap = (uint64_t)s->cfg.patch_ring;
        for (int64_t i = 0; i < n_patches; i++)
            s->patches[(s->patch_published + (uint64_t)i) % pcap] = patches[i];
        const uint64_t seq = s->published;
        ServeReq q{};
        q.solver = (uint32_t)solver;
        q.k = (uint32_t)k;
        q.patch_off = s->patch_published;
        q.n_patches = (uint32_t)n_patches;
        q.seq = seq;
        s->reqs[seq % (uint64_t)s->cfg.request_ring] = q;
        s->pending.push_back(
            Pending{seq, solver, k, s->patch_published, (uint32_t)n_patches, Clock::now()});
        s->patch_published += (uint64_t)n_patches;
        s->pending_words += nw;
        s->published = seq + 1;
        // the entries above are written before the device can see the new head
        __atomic_store_n(s->head, s->published, __ATOMIC_RELEASE);
        *out_id = (int64_t)seq;
    }
    s->cv_work.notify_one();
    return GSS_OK;
}

int gss_server_poll(gss_server *s, int64_t max, int64_t *ids, int64_t *n, int64_t timeout_us) {
    int rc = check(s);
    if (rc) return rc;
    if (!n || max < 0 || (max && !ids)) return fail(GSS_E_INVAL, "bad poll");
    *n = 0;
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->ready.empty() && timeout_us > 0)
        s->cv_done.wait_for(lk, std::chrono::microseconds(timeout_us),
                            [&] { return !s->ready.empty() || s->err_code; });
    while (*n < max && !s->ready.empty()) {
        ids[(*n)++] = s->ready.front();
        s->ready.pop_front();
    }
    if (!*n && s->err_code) return fail(s->err_code, s->err);
    return GSS_OK;
}

int gss_server_result(gss_server *s, int64_t id, gss_serve_result *out) {
    int rc = check(s);
    if (rc) return rc;
    if (!out) return fail(GSS_E_INVAL, "null out");
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->done.find(id);
    if (it == s->done.end()) return fail(GSS_E_STATE, "request not completed (or released)");
    *out = it->second.r;
    return GSS_OK;
}

int gss_server_records(gss_server *s, int64_t id, const gss_serve_hit **out, int64_t *n) {
    int rc = check(s);
    if (rc) return rc;
    if (!out || !n) return fail(GSS_E_INVAL, "null out");
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->done.find(id);
    if (it == s->done.end()) return fail(GSS_E_STATE, "request not completed (or released)");
    Done &d = it->second;
    if (!d.translated) {
        const int rw = d.step->rw;
        d.recs.resize((size_t)d.r.n_records);
        for (int64_t r = 0; r < d.r.n_records; r++) {
            const uint64_t *src = d.step->data() + (d.first + (size_t)r) * rw;
            const uint32_t row = (uint32_t)src[0];
            const uint32_t len = (uint32_t)(src[0] >> 32) & 0xFFFFu;
            gss_serve_hit &h = d.recs[r];
            h.clause = (len < s->ids.size() && row < s->ids[len].size()) ? s->ids[len][row] : -1;
            h.word = (uint32_t)(src[0] >> 48) - (d.step->absolute ? (uint32_t)d.r.first_word : 0u);
            h.pad = 0;
            h.false_mask = rw == 3 ? src[1] : 0;
            h.unit_mask = rw == 3 ? src[2] : 0;
        }
        d.translated = true;
        d.step.reset();  // the step's buffer goes once every request has its records
        d.r.records = d.recs.data();
    }
    *out = d.recs.data();
    *n = d.r.n_records;
    return GSS_OK;
}

int gss_server_release(gss_server *s, int64_t id) {
    int rc = check(s);
    if (rc) return rc;
    std::lock_guard<std::mutex> g(s->mu);
    s->done.erase(id);
    return GSS_OK;
}

int gss_server_stats_get(gss_server *s, gss_server_stats *out) {
    int rc = check(s);
    if (rc) return rc;
    if (!out) return fail(GSS_E_INVAL, "null out");
    std::lock_guard<std::mutex> g(s->mu);
    out->steps = s->steps;
    out->requests_done = s->requests_done;
    out->requests_pending = (int64_t)s->pending.size();
    out->records = s->records;
    out->clauses = s->clauses_admitted;
    out->clauses_refused = s->refused;
    out->literals = s->literals_admitted;
    out->checks = s->checks;
    out->device_ms = s->device_ms;
    out->host_grouped = s->host_grouped;
    out->graph_captures = gss::engine_graph_captures(s->e);
    out->submit_ms_total = s->submit_ms_total;
    out->submit_ms_max = s->submit_ms_max;
    out->wait_ms_max = s->wait_ms_max;
    return GSS_OK;
}

}  // extern "C"
