// Host side of the check engine: device memory for the clause buckets, the
// assignment table and the hit chunks, and the C ABI declared in gss.h.
//
// Layout in HBM
//   clause bucket `len`: uint32 row-major (row r at r*len), capacity doubling;
//       each literal l is stored as l << 7, the byte offset of its table row
//       at 16 words.  Replaces the reference's 32-interleaved per-size arrays
//       (gpuShareLib/Clauses.cu:47-55) because a clause is read by a whole lane
//       group at once here, not by one lane.
//   assignment table: uint64 [2*vars + 2][words]; row l = slots where literal
//       l is NOT false (row 2v: v not false, row 2v+1: v not true), then two
//       sentinel rows.  Replaces per-solver MultiLBool arrays + MultiAgg aggregates
//       (gpuShareLib/Assigs.cuh:39-74): one row holds every solver's slots, so
//       the reference's aggregate prefilter and agg-bit -> solver walk
//       (GpuRunner.cu:92-131) are not needed.
//   hits: chunks of kHitChunk 24-byte records + a fill count per chunk.
//
// Threading: one engine is driven by one host thread at a time.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <thread>
#include <deque>
#include <condition_variable>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "errors.h"
#include "gss.h"
#include "engine_internal.h"
#include "kernels.h"
#include "stage_pool.h"

using namespace gss;

// Scratch of the no-allocation record gather (gss_hits_gather): the record
// offset of every hit chunk the last sweep claimed, then the total.  One copy
// per device, in this library's code object, held until the process ends.
// Host code never uses the array's name as an address: every kernel argument
// takes the per-device address that hipGetSymbolAddress returned for it
// (scratch_address below, called once per device by gss_engine_create with
// that device current).  Every call on one device shares the array, and
// k_chunk_offsets rewrites entries [0, chunks] each time, so nothing in it
// needs clearing; two gathers that overlap on one device race.
static constexpr uint32_t kGatherMaxChunks = 262144;  // 2 MiB + 8 B per device
constexpr int kMaxHostRecordBuffers = 4;  // pinned record buffers per step slot
static __device__ __align__(256) uint64_t scratch[kGatherMaxChunks + 1];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint64_t *g_scratch_ptr[kMaxDevices] = {};

// Clauses of at most this many literals read both rows of each literal in one
// pass (see kernels.hip).
constexpr int kDefaultFuseLen = 4;
// steps staging at least this many bytes split their copies over the stage pool
constexpr size_t kStagePoolMinBytes = 1u << 20;

// Clause literals are stored on the device as row offsets, literal << 7 (the
// byte offset of the literal's row in a 16-word table); see kernels.hip.
constexpr int kLitShift = 7;
constexpr uint32_t kMaxLit = (1u << (32 - kLitShift)) - 1;
// a table of GSS_MAX_VARS variables has 2 * GSS_MAX_VARS + 2 rows (the two
// sentinels included), the last at offset kMaxLit
static_assert(2ull * GSS_MAX_VARS + 1 == kMaxLit, "the exported variable limit follows the row offsets");

thread_local std::string g_err;

int set_err(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIP_TRY(expr)                                                                      \
    do {                                                                                   \
        hipError_t _e = (expr);                                                            \
        if (_e != hipSuccess)                                                              \
            return set_err(_e == hipErrorOutOfMemory ? GSS_E_NOMEM : GSS_E_HIP, "%s: %s", \
                           #expr, hipGetErrorString(_e));                                  \
    } while (0)

// Device address of `scratch` on `device`, which must be current: taken once
// per device, then reused.
int scratch_address(int device, uint64_t **out) {
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (device < 0 || device >= kMaxDevices)
        return set_err(GSS_E_INVAL, "device %d: at most %d devices", device, kMaxDevices);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        const hipError_t he = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (he != hipSuccess || !p)
            return set_err(GSS_E_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", device,
                           hipGetErrorString(he));
        g_scratch_ptr[device] = (uint64_t *)p;
    }
    *out = g_scratch_ptr[device];
    return GSS_OK;
}

struct Bucket {
    uint32_t *d = nullptr;
    uint32_t *tags = nullptr;  // per-row caller tag (tagged DBs), same capacity
    int64_t count = 0;
    int64_t cap = 0;  // rows
};

// Host memory the device copies from or to (or, for staging that kernels
// read directly, reads itself).  Page-locked; a buffer given a page-locked
// limit (one that is only a copy endpoint) falls back to pageable memory when
// an allocation reaches the limit or page-locking fails, as the reference
// does with maxPageLockedMemory (gpuShareLib/CorrespArr.cu:212-235).
template <typename T>
struct Pinned {
    T *p = nullptr;
    size_t cap = 0;
    bool paged = false;
    const int64_t *page_limit = nullptr;  // bytes (< 0: none); null = page-locked only
    int reserve(size_t n) {
        if (n <= cap) return GSS_OK;
        size_t nc = std::max(n, cap * 2);
        release();
        const size_t bytes = nc * sizeof(T);
        const bool lock = !page_limit || *page_limit < 0 || (int64_t)bytes < *page_limit;
        if (lock && hipHostMalloc((void **)&p, bytes, hipHostMallocDefault) == hipSuccess) {
            paged = false;
            cap = nc;
            return GSS_OK;
        }
        p = nullptr;
        if (!page_limit)
            return set_err(GSS_E_NOMEM, "pinned host allocation of %zu bytes failed", bytes);
        p = (T *)std::malloc(bytes);
        if (!p) return set_err(GSS_E_NOMEM, "host allocation of %zu bytes failed", bytes);
        paged = true;
        cap = nc;
        return GSS_OK;
    }
    void release() {
        if (p) {
            if (paged) std::free(p);
            else (void)hipHostFree(p);
        }
        p = nullptr;
        cap = 0;
        paged = false;
    }
    // exchanges the allocations (not the page limit)
    void swap_with(Pinned &o) {
        std::swap(p, o.p);
        std::swap(cap, o.cap);
        std::swap(paged, o.paged);
    }
    ~Pinned() { release(); }
};

// Allocates and frees pinned host record buffers on a thread of its own, so
// that the thread waiting for a step does not stall for the milliseconds a
// large pinned allocation (or its release) takes: a buffer's replacement is
// allocated ahead of the wait that needs it.
class HostPrefetch {
  public:
    explicit HostPrefetch(int device) : device_(device), th_([this] { run(); }) {}
    ~HostPrefetch() {
        {
            std::lock_guard<std::mutex> g(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        th_.join();
    }
    // dst gets at least n records; *state goes from 1 to 2 when it is ready
    // (0 if the allocation failed)
    void allocate(Pinned<gss_hit> *dst, std::atomic<int> *state, size_t n) {
        state->store(1, std::memory_order_release);
        {
            std::lock_guard<std::mutex> g(mu_);
            jobs_.push_back({dst, state, n, nullptr, 0, false});
        }
        cv_.notify_all();
    }
    // takes over buf's allocation and releases it
    void release(Pinned<gss_hit> &buf) {
        if (!buf.p) return;
        {
            std::lock_guard<std::mutex> g(mu_);
            jobs_.push_back({nullptr, nullptr, 0, buf.p, buf.cap, buf.paged});
        }
        buf.p = nullptr;
        buf.cap = 0;
        buf.paged = false;
        cv_.notify_all();
    }
    void wait(const std::atomic<int> *state) {
        std::unique_lock<std::mutex> g(mu_);
        done_.wait(g, [&] { return state->load(std::memory_order_acquire) != 1; });
    }

  private:
    struct Job {
        Pinned<gss_hit> *dst;
        std::atomic<int> *state;
        size_t n;
        gss_hit *old;  // release job: the allocation to free
        size_t old_cap;
        bool old_paged;
    };
    void run() {
        (void)hipSetDevice(device_);
        for (;;) {
            Job j;
            {
                std::unique_lock<std::mutex> g(mu_);
                cv_.wait(g, [&] { return stop_ || !jobs_.empty(); });
                if (jobs_.empty()) return;  // stopping, nothing left
                j = jobs_.front();
                jobs_.pop_front();
            }
            const auto t0 = std::chrono::steady_clock::now();
            if (j.dst) {
                const int rc = j.dst->reserve(j.n);
                gss::log_stall("record buffer allocated ahead",
                          std::chrono::duration<double, std::milli>(
                              std::chrono::steady_clock::now() - t0)
                              .count());
                {
                    std::lock_guard<std::mutex> g(mu_);
                    j.state->store(rc == GSS_OK ? 2 : 0, std::memory_order_release);
                }
                done_.notify_all();
            } else if (j.old) {
                if (j.old_paged) std::free(j.old);
                else (void)hipHostFree(j.old);
                gss::log_stall("record buffer freed",
                          std::chrono::duration<double, std::milli>(
                              std::chrono::steady_clock::now() - t0)
                              .count());
            }
        }
    }
    int device_;
    std::mutex mu_;
    std::condition_variable cv_, done_;
    std::deque<Job> jobs_;
    bool stop_ = false;
    std::thread th_;
};

template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t cap = 0;
    int reserve(size_t n, bool keep = false, hipStream_t s = nullptr) {
        if (n <= cap) return GSS_OK;
        size_t nc = std::max(n, cap + cap / 2);
        T *np = nullptr;
        hipError_t e = hipMalloc((void **)&np, nc * sizeof(T));
        if (e != hipSuccess)
            return set_err(GSS_E_NOMEM, "device allocation of %zu bytes failed: %s",
                           nc * sizeof(T), hipGetErrorString(e));
        if (p) {
            if (keep && cap) {
                HIP_TRY(hipMemcpyAsync(np, p, cap * sizeof(T), hipMemcpyDeviceToDevice, s));
            }
            // the old buffer may still be read by queued work on any stream
            HIP_TRY(hipDeviceSynchronize());
            (void)hipFree(p);
        }
        p = np;
        cap = nc;
        return GSS_OK;
    }
    // Stream-ordered growth (contents not kept): no host synchronisation.
    int reserve_async(size_t n, hipStream_t s) {
        if (n <= cap) return GSS_OK;
        size_t nc = std::max(n, cap * 2);
        T *np = nullptr;
        hipError_t e = hipMallocAsync((void **)&np, nc * sizeof(T), s);
        if (e != hipSuccess)
            return set_err(GSS_E_NOMEM, "device allocation of %zu bytes failed: %s",
                           nc * sizeof(T), hipGetErrorString(e));
        if (p) HIP_TRY(hipFreeAsync(p, s));
        p = np;
        cap = nc;
        async = true;
        return GSS_OK;
    }
    void release() {
        if (p) (void)(async ? hipFreeAsync(p, nullptr) : hipFree(p));
        p = nullptr;
        cap = 0;
    }
    bool async = false;
};

// One of the two ping-pong slots of batched steps: pinned staging the host
// fills, the device copies the step's kernels read, this slot's hit buffer,
// and the graph captured over exactly these buffers.
struct StepSlot {
    Pinned<StepHeader> h_hdr;
    Pinned<BucketDesc> h_desc;
    Pinned<ClauseSeg> h_cls;
    Pinned<uint32_t> h_lits;
    Pinned<gss_var_update> h_ups;
    Pinned<gss_var_patch> h_patches;
    Pinned<gss_base_update> h_bups;
    Pinned<gss_word_load> h_loads;
    Pinned<uint32_t> h_out;  // [0] = chunks claimed, [1 + c] = fill of chunk c
    DevBuf<StepHeader> d_hdr;
    DevBuf<BucketDesc> d_desc;
    DevBuf<ClauseSeg> d_cls;
    DevBuf<uint32_t> d_lits;
    DevBuf<gss_var_update> d_ups;
    DevBuf<gss_var_patch> d_patches;
    DevBuf<gss_base_update> d_bups;
    DevBuf<gss_word_load> d_loads;
    DevBuf<uint64_t> d_hits;     // chunk_cap chunks of kHitChunk records of rec_words words
    DevBuf<uint32_t> d_fills;
    DevBuf<gss_hit> d_dense;     // compacted records of the last waited step
    DevBuf<uint64_t> d_offs;     // record offset of each chunk (the graph's compaction)
    uint64_t dense_want = 0;     // records the graph's compaction should hold
    uint64_t step_dense = 0;     // records the last submitted step's graph compacts into
                                 // d_dense (0 = none)
    // The records of a waited step go to a pinned host buffer on the slot's
    // own stream as soon as gss_step_wait has seen the step end (one DMA copy
    // when the graph compacted them); the slot's next graph waits for that
    // copy on the engine stream, so the records stay valid until the slot's
    // next step is waited and the host can stage that step while they travel.
    // With h_n > 1 buffers (serving engines) the waits rotate through them,
    // and a step's records stay in place for h_n waits of the slot.
    Pinned<gss_hit> h_ring[kMaxHostRecordBuffers];
    // replacements of h_ring[i] allocated ahead by the engine's HostPrefetch
    // (state 0: none, 1: being allocated, 2: ready)
    Pinned<gss_hit> h_spare[kMaxHostRecordBuffers];
    std::atomic<int> spare_state[kMaxHostRecordBuffers] = {};
    int h_at = 0, h_n = 1;
    Pinned<gss_hit> &h_hits() { return h_ring[h_at]; }
    hipStream_t copy = nullptr;
    hipEvent_t ev_copied = nullptr;
    bool copy_queued = false;
    int64_t rec_total = 0;  // records of the slot's last waited step
    bool host_ready = false;     // h_hits() holds the last waited step's records
    Pinned<ScatterSeg> h_cseg;   // compaction segments (one per non-empty chunk)
    DevBuf<ScatterSeg> d_cseg;
    uint32_t *d_alloc = nullptr;
    uint32_t chunk_cap = 0;
    // chunks that hold every record the step could make; a buffer sized by
    // the records seen (chunk_cap < worst_chunks) may overflow, and the step
    // is then recovered (its sweep re-run) before the table moves on
    uint64_t worst_chunks = 0;
    bool settled = false;       // the step in flight is known to have kept its records
    bool records_lost = false;  // its recovery could not grow the buffer
    // launch settings of the last submitted step's graph, for a recovery's
    // re-run of its tail
    int g_blocks = 0, g_tc = 0, g_fuse = 0, g_unroll = 0;
    bool g_serve = false, g_screen = false;
    bool g_agg_full = false;  // screened: the aggregate is rebuilt from the table
    // Graphs captured over this slot's buffers, one per key (everything baked
    // in); a caller alternating between step forms (screened or not, rebuilt
    // aggregate or refreshed) replays them without capturing again.  The
    // least recently used one goes when a fifth is needed.
    struct StepGraph {
        std::vector<const void *> key;
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        int blocks = 0, tc = 0, fuse = 0, unroll = 0;
        bool serve = false, screen = false, agg_full = false;
        uint64_t dense = 0;  // records it compacts into d_dense (0 = none)
        uint64_t used = 0;   // submit count at its last use
    };
    static constexpr size_t kGraphsPerSlot = 4;
    std::vector<StepGraph> graphs;
    int cur = -1;  // graph of the slot's last submitted step
    uint64_t submits = 0;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    bool busy = false;
    bool serve = false;    // the slot's last step admitted requests from the serving ring
    // the slot's graph holds the base-plane nodes (base updates, word loads)
    // and the patch node once a step of the slot needed them (then always:
    // a caller alternating between step kinds does not re-capture each time)
    bool g_bases = false, g_patches = false;
    Pinned<ServeReport> h_report;  // what that step admitted (serving steps)
    // serving steps: records made dense grouped by assignment word, and where
    // each word's records start (h_group[w], [64] = all, [65] = 1 if grouped)
    DevBuf<uint32_t> d_wcount;     // 64 counts + 64 cursors
    Pinned<uint64_t> h_group;
    bool results = false;  // h_out / d_hits hold the results of the last waited step
    int rec_words = 3;     // record form of the slot's last submitted step (3: gss_hit,
                           // 1: gss_hit_id, 2: gss_hit_clause)
    int res_words = 3;     // the same, of its last waited step (the results)
};

}  // namespace

int gss::record_error(int code, const char *msg) {
    g_err = msg ? msg : "";
    return code;
}

void gss::log_stall(const char *what, double ms) {
    static const double limit = [] {
        const char *v = std::getenv("GPUSHARESAT_LOG_STALLS_MS");
        return v ? std::atof(v) : -1.0;
    }();
    static const auto t0 = std::chrono::steady_clock::now();
    if (limit < 0 || ms < limit) return;
    const double at =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::fprintf(stderr, "[gpusharesat] %.3f s: %s took %.1f ms\n", at, what, ms);
}

struct gss_engine {
    int device = 0;
    int cu = 0;
    hipStream_t stream = nullptr;
    hipStream_t capture_stream = nullptr;  // records step graphs (never executes work)
    hipEvent_t ev0 = nullptr, ev1 = nullptr;

    std::vector<Bucket> buckets;  // index = clause length
    bool tagged = false;          // every row carries a uint32 tag
    // bumped whenever existing rows move or vanish (compaction, bucket
    // replacement, clear); appends leave existing rows in place
    uint64_t rows_version = 0, sweep_rows_version = 0;
    // bumped by every call that changes the clause DB or the assignment
    // table; a plain sweep whose hit buffer overflowed is re-run only if
    // nothing changed since it was queued (else the re-run would check a
    // different DB or table than the caller asked for)
    uint64_t state_version = 0, sweep_state_version = 0;
    int64_t total_clauses = 0, total_lits = 0;
    // Cap on the device bytes of the clause buckets (their capacities), from
    // GPUSHARESAT_DB_LIMIT_BYTES at creation; 0 = only the device's memory.
    // Growth past it fails with GSS_E_NOMEM exactly as a full device would.
    int64_t db_limit = 0;
    // Largest literal ever stored (running max, never lowered); a sweep is
    // refused unless the table has a row for it, so the kernel's table reads
    // stay in bounds.
    int64_t max_lit = -1;

    // assignment table
    DevBuf<uint64_t> table;
    int64_t vars = 0;
    int words = 0;
    // solver base planes (gss_bases_reset): [nbases][2][bases_vw] bits
    DevBuf<uint64_t> bases;
    int nbases = 0;
    int64_t bases_vw = 0;

    // sweep state
    DevBuf<BucketDesc> d_desc;
    Pinned<BucketDesc> h_desc;
    uint64_t *d_valid = nullptr;
    Pinned<uint64_t> h_valid;
    // plain sweep: what d_desc / d_valid hold (uploads are skipped when a
    // sweep's descriptors and slot mask are unchanged), and the pinned copy of
    // the claim counter and record count that k_count_records writes after
    // the sweep
    std::vector<BucketDesc> desc_new, desc_up;
    const BucketDesc *desc_up_at = nullptr;
    std::vector<uint64_t> valid_up;
    Pinned<uint32_t> h_count;  // [0] = chunks claimed, [1], [2] = records (low, high word)
    uint64_t *gather_scratch = nullptr;  // this device's address of `scratch`
    // host buffers that are only copy endpoints of at least this many bytes
    // are pageable (< 0: no limit); see Pinned
    int64_t page_limit = -1;
    DevBuf<gss_hit> d_hits;
    DevBuf<uint32_t> d_fills;
    uint32_t *d_alloc = nullptr;
    uint32_t chunk_cap = 0;
    int nbuckets = 0, ntasks = 0, task_clauses = 0, blocks = 0;
    bool pending = false, waited = false, have_sweep = false;
    int64_t last_hits = 0;
    uint32_t last_chunks = 0;
    float last_ms = 0.f;
    double total_ms = 0.0;
    int64_t nsweeps = 0;
    int occ[7][9][2][4] = {};  // [log2 words][unroll][fused][sweep mode] resident blocks per CU

    // tuning
    int tune_waves_per_cu = 0;
    int tune_task_clauses = 0;
    int tune_max_chunks = 0;
    int tune_unroll = 0;
    int tune_fuse_len = 0;  // 0 = default, -1 = no fused schedule
    int tune_schedule = 0;  // 0 = default (long clauses first), 1 = long first, 2 = short first

    // staging
    Pinned<uint32_t> h_stage;
    DevBuf<uint32_t> d_stage;
    Pinned<ScatterSeg> h_segs;
    DevBuf<ScatterSeg> d_segs;
    Pinned<gss_var_update> h_ups;
    DevBuf<gss_var_update> d_ups;
    int8_t *d_lastpos = nullptr;
    Pinned<int8_t> h_lastpos;
    int64_t applied = 0;  // updates of the last gss_assig_apply, still in d_ups
    DevBuf<uint64_t> d_offsets;  // record offset of each chunk (dense copies of a sweep)
    DevBuf<gss_hit> d_dense;
    Pinned<gss_hit> h_hits;

    // tagged DB compaction
    DevBuf<CompactRef> d_cref;
    DevBuf<uint32_t> d_flags, d_pos, d_kept;
    DevBuf<uint8_t> d_scan_temp, d_keep;
    Pinned<uint32_t> h_kept;

    // serving ring (gss_server): requests admitted on the device by serving steps
    bool serve_on = false;
    ServeRing serve{};
    DevBuf<uint64_t> serve_tail;
    DevBuf<ServeAdmitted> serve_adm;
    DevBuf<uint32_t> serve_cnt;
    // aggregate entries screened ahead of the exact check (SweepArgs::agg), by
    // serving steps (from the requests' bases; GPUSHARESAT_SERVE_PREFILTER=0
    // leaves them out, same records) and by batched steps submitted with
    // GSS_STEP_SCREEN (from the table)
    bool serve_prefilter = true;
    int serve_vote = 2;  // entries per screening vote (GPUSHARESAT_SERVE_VOTE: 1, 2, 4 or 8)
    DevBuf<uint8_t> serve_agg;         // 2 * vars + 2 bytes, padded to 4
    DevBuf<uint64_t> serve_agg_words;  // [8]
    // state_version right after the last screened batched step was submitted:
    // while it still holds, serve_agg is that step's aggregate kept current
    // (refreshed from its updates after the roll forward), so the next
    // screened step only refreshes what its own updates change
    uint64_t table_agg_at = 0;
    DevBuf<gss_var_patch> serve_dpatches;  // the step's admitted patches (ring capacity)

    // batched steps (gss_step_*)
    StepSlot slots[2];
    int next_slot = 0;
    std::vector<int> inflight;  // slot indices, oldest first
    int last_step_slot = -1;
    int64_t last_step_hits = 0;
    uint32_t last_step_chunks = 0;
    float last_step_ms = 0.f;
    std::atomic<int64_t> graph_captures{0};  // step graphs captured (read by other threads)
    // Step hit buffers are sized by the records steps have made: twice the
    // most chunks a step claimed (at least a floor), never beyond the worst
    // case.  A step that overflows its buffer sets d_halt (device, zero while
    // running); later steps in flight then stage nothing until its wait (or
    // a plain call) re-runs its sweep with room and replays them.
    uint32_t *d_halt = nullptr;
    uint64_t step_hit_want = 0;
    std::atomic<int64_t> step_recoveries{0};
    std::atomic<int64_t> step_hit_chunks{0};  // chunks the two slots' hit buffers hold
    // pinned host record buffers grown in gss_step_wait, and the time taken
    std::atomic<int64_t> host_grows{0}, host_grow_us{0}, host_grow_us_max{0};
    uint64_t records_hw = 0;  // the most records a waited step has made
    std::unique_ptr<HostPrefetch> prefetch;  // created at the first record-buffer growth
    // host threads for the staging copies of large steps
    // (GPUSHARESAT_STAGE_THREADS, default 4; 0 or 1 = the calling thread only)
    int stage_threads = 4;
    std::unique_ptr<gss::StagePool> stage_pool;
};

namespace {

int check_engine(gss_engine *e) {
    if (!e) return set_err(GSS_E_INVAL, "null engine");
    return hipSetDevice(e->device) == hipSuccess ? GSS_OK
                                                 : set_err(GSS_E_HIP, "hipSetDevice failed");
}

}  // namespace

static int settle_steps(gss_engine *e);

namespace {

// Fuse length the kernel runs with: <= 0 selects the variant without the
// fused schedule.
int fuse_of(const gss_engine *e) {
    return e->tune_fuse_len > 0 ? e->tune_fuse_len : e->tune_fuse_len < 0 ? 0 : kDefaultFuseLen;
}

// Resident 256-thread blocks per CU of the sweep variant the engine runs.
int resident_blocks(gss_engine *e, int mode = 0) {
    int l = 0;
    while ((1 << l) < e->words) l++;
    const int ui = std::min(e->tune_unroll, 8);
    const int fi = fuse_of(e) > 0;
    int &o = e->occ[l][ui][fi][mode];
    if (!o) o = sweep_occupancy(e->words, e->tune_unroll, fi, mode);
    return o;
}

int log2_exact(int w) {
    int l = 0;
    while ((1 << l) < w) l++;
    return (1 << l) == w ? l : -1;
}

// With shrink, an array more than twice the need is replaced by one of the
// need (a bucket being replaced, count already 0): a reduced DB gives its
// memory back.  The new array is allocated before the old one is released.
int ensure_bucket_cap(gss_engine *e, int len, int64_t need, bool shrink = false) {
    if ((int)e->buckets.size() <= len) e->buckets.resize(len + 1);
    Bucket &b = e->buckets[len];
    const bool smaller = shrink && b.d && std::max<int64_t>(need, 256) < b.cap / 2;
    if (need <= b.cap && !smaller) return GSS_OK;
    int64_t nc = smaller ? std::max<int64_t>(need, 256)
                         : std::max<int64_t>(std::max<int64_t>(need, b.cap * 2), 256);
    if (e->db_limit > 0) {
        const int64_t per_row = (int64_t)len * 4 + (e->tagged ? 4 : 0);
        int64_t used = 0;
        for (size_t l = 1; l < e->buckets.size(); l++)
            used += e->buckets[l].cap * ((int64_t)l * 4 + (e->tagged ? 4 : 0));
        const int64_t room = e->db_limit - (used - b.cap * per_row);
        if (nc * per_row > room) nc = need;  // no headroom rather than no room
        if (nc * per_row > room)
            return set_err(GSS_E_NOMEM, "bucket %d: %lld rows exceed the clause DB limit of %lld bytes",
                           len, (long long)nc, (long long)e->db_limit);
    }
    // Stream-ordered: the new array is allocated, filled and the old one
    // released in the engine stream's order, after every queued sweep or step
    // that still reads the old one, so growth never stalls the host.
    const auto t0 = std::chrono::steady_clock::now();
    uint32_t *np = nullptr, *nt = nullptr;
    hipError_t he = hipMallocAsync((void **)&np, (size_t)nc * len * sizeof(uint32_t), e->stream);
    if (he == hipSuccess && e->tagged) {
        he = hipMallocAsync((void **)&nt, (size_t)nc * sizeof(uint32_t), e->stream);
        if (he != hipSuccess) (void)hipFreeAsync(np, e->stream);
    }
    if (he != hipSuccess)
        return set_err(GSS_E_NOMEM, "bucket %d: device allocation of %lld rows failed: %s", len,
                       (long long)nc, hipGetErrorString(he));
    if (b.d) {
        if (b.count) {
            HIP_TRY(hipMemcpyAsync(np, b.d, (size_t)b.count * len * sizeof(uint32_t),
                                   hipMemcpyDeviceToDevice, e->stream));
            if (nt && b.tags)
                HIP_TRY(hipMemcpyAsync(nt, b.tags, (size_t)b.count * sizeof(uint32_t),
                                       hipMemcpyDeviceToDevice, e->stream));
        }
        HIP_TRY(hipFreeAsync(b.d, e->stream));
    }
    if (b.tags) HIP_TRY(hipFreeAsync(b.tags, e->stream));
    b.d = np;
    b.tags = nt;
    b.cap = nc;
    gss::log_stall("clause bucket growth",
                   std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
                       .count());
    return GSS_OK;
}

// Launches the sweep kernel for the current descriptors (already uploaded).
int launch_current_sweep(gss_engine *e) {
    HIP_TRY(launch_reset_u32(e->d_alloc, 1, e->stream));
    HIP_TRY(hipEventRecord(e->ev0, e->stream));
    if (e->ntasks > 0) {
        SweepArgs a{};
        a.dyn = nullptr;  // counts passed by value
        a.buckets = e->d_desc.p;
        a.nbuckets = e->nbuckets;
        a.ntasks = e->ntasks;
        a.task_clauses = e->task_clauses;
        a.table = e->table.p;
        a.sentinel = (uint32_t)(2 * e->vars) << kLitShift;
        a.valid = e->d_valid;
        a.hits = e->d_hits.p;
        a.fills = e->d_fills.p;
        a.alloc = e->d_alloc;
        a.chunk_cap = e->chunk_cap;
        a.fuse_len = fuse_of(e);
        a.rec_words = 3;
        HIP_TRY(launch_sweep(a, e->words, e->tune_unroll, e->blocks, e->stream));
    }
    HIP_TRY(hipEventRecord(e->ev1, e->stream));
    // claim counter and record count to pinned memory: the wait reads them there
    HIP_TRY(launch_count_records(e->d_alloc, e->d_fills.p, e->chunk_cap, e->h_count.p, e->stream));
    return GSS_OK;
}

int reserve_chunks(gss_engine *e, uint64_t chunks) {
    if (chunks > 0xFFFFFFF0ull) return set_err(GSS_E_NOMEM, "hit buffer too large");
    int rc;
    if ((rc = e->d_hits.reserve((size_t)chunks * kHitChunk))) return rc;
    if ((rc = e->d_fills.reserve((size_t)chunks))) return rc;
    if ((rc = e->h_count.reserve(4))) return rc;
    e->chunk_cap = (uint32_t)std::min<size_t>(e->d_fills.cap, e->d_hits.cap / kHitChunk);
    return GSS_OK;
}

}  // namespace

extern "C" {

const char *gss_last_error(void) { return g_err.c_str(); }

int gss_engine_create(int device, gss_engine **out) {
    if (!out) return set_err(GSS_E_INVAL, "null out");
    *out = nullptr;
    int ndev = 0;
    HIP_TRY(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev)
        return set_err(GSS_E_INVAL, "device %d out of range (%d devices)", device, ndev);
    HIP_TRY(hipSetDevice(device));
    gss_engine *e = new (std::nothrow) gss_engine();
    if (!e) return set_err(GSS_E_NOMEM, "engine allocation failed");
    e->device = device;
    if (const char *lim = std::getenv("GPUSHARESAT_DB_LIMIT_BYTES"))
        e->db_limit = std::max<long long>(0, std::atoll(lim));
    if (const char *st = std::getenv("GPUSHARESAT_STAGE_THREADS"))
        e->stage_threads = std::min(64, std::max(0, std::atoi(st)));
    if (const char *pf = std::getenv("GPUSHARESAT_SERVE_PREFILTER"))
        e->serve_prefilter = std::atoi(pf) != 0;
    if (const char *sv = std::getenv("GPUSHARESAT_SERVE_VOTE")) {
        const int v = std::atoi(sv);
        if (v == 1 || v == 2 || v == 4 || v == 8) e->serve_vote = v;
    }
    int rc = GSS_OK;
    auto fail = [&](int code) {
        gss_engine_destroy(e);
        return code;
    };
    if (hipDeviceGetAttribute(&e->cu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
        return fail(set_err(GSS_E_HIP, "cannot query CU count"));
    if ((rc = scratch_address(device, &e->gather_scratch))) return fail(rc);
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess)
        return fail(set_err(GSS_E_HIP, "stream creation failed"));
    if (hipEventCreate(&e->ev0) != hipSuccess || hipEventCreate(&e->ev1) != hipSuccess)
        return fail(set_err(GSS_E_HIP, "event creation failed"));
    {
        // memory released in stream order stays in the device pool for the
        // next growth instead of going back to the driver
        hipMemPool_t pool = nullptr;
        uint64_t keep = UINT64_MAX;
        if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess && pool)
            (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
    }
    if (hipMalloc((void **)&e->d_valid, GSS_MAX_WORDS * sizeof(uint64_t)) != hipSuccess ||
        hipMalloc((void **)&e->d_alloc, 64) != hipSuccess ||
        hipMalloc((void **)&e->d_halt, 64) != hipSuccess ||
        hipMalloc((void **)&e->d_lastpos, GSS_MAX_WORDS) != hipSuccess)
        return fail(set_err(GSS_E_NOMEM, "device allocation failed"));
    if (hipMemset(e->d_halt, 0, 64) != hipSuccess)
        return fail(set_err(GSS_E_HIP, "device memset failed"));
    if ((rc = e->h_valid.reserve(GSS_MAX_WORDS)) || (rc = e->h_lastpos.reserve(GSS_MAX_WORDS)))
        return fail(rc);
    // copy endpoints only: staging of the plain path, the steps' records
    e->h_stage.page_limit = &e->page_limit;
    e->h_ups.page_limit = &e->page_limit;
    for (StepSlot &sl : e->slots)
        for (auto &h : sl.h_ring) h.page_limit = &e->page_limit;
    for (StepSlot &sl : e->slots)
        for (auto &h : sl.h_spare) h.page_limit = &e->page_limit;
    if ((rc = gss_assig_reset(e, 0, 1))) return fail(rc);
    *out = e;
    return GSS_OK;
}

int gss_engine_destroy(gss_engine *e) {
    if (!e) return GSS_OK;
    (void)hipSetDevice(e->device);
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    e->prefetch.reset();  // its thread may be filling a slot's spare buffer
    for (StepSlot &sl : e->slots) {
        if (sl.copy) (void)hipStreamSynchronize(sl.copy);
        for (StepSlot::StepGraph &g : sl.graphs) {
            if (g.exec) (void)hipGraphExecDestroy(g.exec);
            if (g.graph) (void)hipGraphDestroy(g.graph);
        }
        if (sl.t0) (void)hipEventDestroy(sl.t0);
        if (sl.t1) (void)hipEventDestroy(sl.t1);
        if (sl.ev_copied) (void)hipEventDestroy(sl.ev_copied);
        if (sl.copy) (void)hipStreamDestroy(sl.copy);
        if (sl.d_alloc) (void)hipFree(sl.d_alloc);
        sl.d_hdr.release();
        sl.d_desc.release();
        sl.d_cls.release();
        sl.d_lits.release();
        sl.d_ups.release();
        sl.d_patches.release();
        sl.d_bups.release();
        sl.d_loads.release();
        sl.d_hits.release();
        sl.d_fills.release();
        sl.d_dense.release();
        sl.d_offs.release();
        sl.d_wcount.release();
        sl.d_cseg.release();
    }
    for (Bucket &b : e->buckets) {
        if (b.d) (void)hipFreeAsync(b.d, e->stream);
        if (b.tags) (void)hipFreeAsync(b.tags, e->stream);
    }
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    e->table.release();
    e->bases.release();
    e->d_desc.release();
    e->d_hits.release();
    e->d_fills.release();
    e->d_stage.release();
    e->d_segs.release();
    e->d_ups.release();
    e->d_offsets.release();
    e->d_dense.release();
    e->serve_tail.release();
    e->serve_adm.release();
    e->serve_cnt.release();
    e->serve_agg.release();
    e->serve_agg_words.release();
    e->serve_dpatches.release();
    if (e->d_valid) (void)hipFree(e->d_valid);
    if (e->d_alloc) (void)hipFree(e->d_alloc);
    if (e->d_halt) (void)hipFree(e->d_halt);
    if (e->d_lastpos) (void)hipFree(e->d_lastpos);
    if (e->ev0) (void)hipEventDestroy(e->ev0);
    if (e->ev1) (void)hipEventDestroy(e->ev1);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    if (e->capture_stream) (void)hipStreamDestroy(e->capture_stream);
    delete e;
    return GSS_OK;
}

int gss_engine_cu_count(gss_engine *e, int *out) {
    int rc = check_engine(e);
    if (rc) return rc;
    if (!out) return set_err(GSS_E_INVAL, "null out");
    *out = e->cu;
    return GSS_OK;
}

int gss_engine_sync(gss_engine *e) {
    int rc = check_engine(e);
    if (rc) return rc;
    if ((rc = settle_steps(e))) return rc;
    HIP_TRY(hipStreamSynchronize(e->stream));
    return GSS_OK;
}

int gss_engine_set_page_locked_limit(gss_engine *e, int64_t bytes) {
    int rc = check_engine(e);
    if (rc) return rc;
    e->page_limit = bytes < 0 ? -1 : bytes;
    return GSS_OK;
}

int gss_engine_mem_info(gss_engine *e, size_t *free_bytes, size_t *total_bytes) {
    int rc = check_engine(e);
    if (rc) return rc;
    size_t f = 0, t = 0;
    HIP_TRY(hipMemGetInfo(&f, &t));
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
    return GSS_OK;
}

/* ---- clause DB ------------------------------------------------------------ */

static int db_append(gss_engine *e, int64_t n, const uint32_t *lits, const int32_t *lens,
                     const uint32_t *tags, uint32_t *out_rows) {
    int rc = check_engine(e);
    if (rc) return rc;
    if ((rc = settle_steps(e))) return rc;
    if (n < 0 || (n > 0 && (!lits || !lens))) return set_err(GSS_E_INVAL, "bad clause batch");
    if (e->tagged && n > 0 && !tags) return set_err(GSS_E_INVAL, "tagged DB: tags required");
    if (n == 0) return GSS_OK;
    // Group by length (stable), so that each bucket's new rows are one
    // contiguous segment of the staging buffer.
    int maxlen = 0;
    std::vector<int64_t> start(n);
    int64_t pos = 0;
    for (int64_t i = 0; i < n; i++) {
        if (lens[i] < 1 || lens[i] > GSS_MAX_CLAUSE_LEN)
            return set_err(GSS_E_INVAL, "clause %lld: length %d outside [1, %d]", (long long)i,
                           lens[i], GSS_MAX_CLAUSE_LEN);
        maxlen = std::max(maxlen, (int)lens[i]);
        start[i] = pos;
        pos += lens[i];
    }
    const int64_t total = pos;
    uint32_t mx = 0;
    for (int64_t i = 0; i < total; i++) mx = std::max(mx, lits[i]);
    if (mx > kMaxLit) return set_err(GSS_E_INVAL, "literal %u exceeds %u", mx, kMaxLit);
    std::vector<int64_t> per(maxlen + 1, 0);
    for (int64_t i = 0; i < n; i++) per[lens[i]]++;
    if ((int)e->buckets.size() <= maxlen) e->buckets.resize(maxlen + 1);
    for (int len = 1; len <= maxlen; len++)
        if (per[len] && (rc = ensure_bucket_cap(e, len, e->buckets[len].count + per[len])))
            return rc;
    // staging offsets (in literals) of each bucket's segment
    std::vector<int64_t> seg_off(maxlen + 1, 0), seg_fill(maxlen + 1, 0);
    int64_t acc = 0;
    int nseg = 0;
    for (int len = 1; len <= maxlen; len++) {
        seg_off[len] = acc;
        acc += per[len] * len;
        if (per[len]) nseg++;
    }
    // staging: literals, then (tagged DBs) one tag per clause in the same
    // bucket order; one segment per bucket for each
    const int64_t ntag = e->tagged ? n : 0;
    const int nsegs_all = e->tagged ? 2 * nseg : nseg;
    if ((rc = e->h_stage.reserve((size_t)(total + ntag))) ||
        (rc = e->d_stage.reserve((size_t)(total + ntag))) ||
        (rc = e->h_segs.reserve(nsegs_all)) || (rc = e->d_segs.reserve(nsegs_all)))
        return rc;
    // any previous async user of the staging buffers must be done
    HIP_TRY(hipStreamSynchronize(e->stream));
    std::vector<int64_t> tag_off(maxlen + 1, 0);
    if (e->tagged) {
        int64_t ta = total;
        for (int len = 1; len <= maxlen; len++) {
            tag_off[len] = ta;
            ta += per[len];
        }
    }
    for (int64_t i = 0; i < n; i++) {
        const int len = lens[i];
        const int64_t k = seg_fill[len]++;
        std::memcpy(e->h_stage.p + seg_off[len] + k * len, lits + start[i], len * sizeof(uint32_t));
        if (e->tagged) e->h_stage.p[tag_off[len] + k] = tags[i];
        if (out_rows) out_rows[i] = (uint32_t)(e->buckets[len].count + k);
    }
    int s = 0;
    for (int len = 1; len <= maxlen; len++) {
        if (!per[len]) continue;
        Bucket &b = e->buckets[len];
        ScatterSeg &sg = e->h_segs.p[s];
        sg.dst = b.d + (size_t)b.count * len;
        sg.src = (uint64_t)seg_off[len];
        sg.count = (uint64_t)per[len] * len;
        if (e->tagged) {
            ScatterSeg &tg = e->h_segs.p[nseg + s];
            tg.dst = b.tags + b.count;
            tg.src = (uint64_t)tag_off[len];
            tg.count = (uint64_t)per[len];
        }
        s++;
        b.count += per[len];
    }
    HIP_TRY(hipMemcpyAsync(e->d_stage.p, e->h_stage.p, (size_t)(total + ntag) * sizeof(uint32_t),
                           hipMemcpyHostToDevice, e->stream));
    HIP_TRY(hipMemcpyAsync(e->d_segs.p, e->h_segs.p, (size_t)nsegs_all * sizeof(ScatterSeg),
                           hipMemcpyHostToDevice, e->stream));
    HIP_TRY(launch_scatter(e->d_segs.p, nseg, e->d_stage.p, kLitShift, e->stream));
    if (e->tagged)
        HIP_TRY(launch_scatter(e->d_segs.p + nseg, nseg, e->d_stage.p, 0, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    e->total_clauses += n;
    e->total_lits += total;
    e->max_lit = std::max<int64_t>(e->max_lit, mx);
    e->state_version++;
    return GSS_OK;
}

int gss_db_append(gss_engine *e, int64_t n, const uint32_t *lits, const int32_t *lens,
                  uint32_t *out_rows) {
    return db_append(e, n, lits, lens, nullptr, out_rows);
}

int gss_db_append_tagged(gss_engine *e, int64_t n, const uint32_t *lits, const int32_t *lens,
                         const uint32_t *tags, uint32_t *out_rows) {
    if (e && !e->tagged) return set_err(GSS_E_STATE, "the DB is not tagged");
    return db_append(e, n, lits, lens, tags, out_rows);
}

int gss_db_enable_tags(gss_engine *e) {
    int rc = check_engine(e);
    if (rc) return rc;
    if ((rc = settle_steps(e))) return rc;
    if (e->tagged) return GSS_OK;
    if (e->total_clauses) return set_err(GSS_E_STATE, "tags must be enabled on an empty DB");
    for (Bucket &b : e->buckets)
        if (b.cap && !b.tags && hipMalloc((void **)&b.tags, (size_t)b.cap * 4) != hipSuccess)
            return set_err(GSS_E_NOMEM, "tag allocation failed");
    e->tagged = true;
    return GSS_OK;
}

int gss_db_read_tags(gss_engine *e, int len, int64_t row0, int64_t n, uint32_t *out) {
    int rc = check_engine(e);
    if (rc) return rc;
    if ((rc = settle_steps(e))) return rc;
    if (!e->tagged) return set_err(GSS_E_STATE, "the DB is not tagged");
    if (len < 1 || len >= (int)e->buckets.size() || row0 < 0 || n < 0 ||
        row0 + n > e->buckets[len].count || (n && !out))
        return set_err(GSS_E_INVAL, "bad tag read");
    if (n)
        HIP_TRY(hipMemcpy(out, e->buckets[len].tags + row0, (size_t)n * 4,
                          hipMemcpyDeviceToHost));
    return GSS_OK;
}

}  // extern "C"

// Keeps the rows of a tagged DB whose tag t has d_keep[t] != 0 (tags >= ntags
// are kept), in order, on the device: flags -> exclusive scan -> per-bucket
// kept counts -> moves into fresh arrays.
int gss::db_compact_device(gss_engine *e, const uint8_t *d_keep, int64_t ntags) {
    int rc = check_engine(e);
    if (rc) return rc;
    if (!e->tagged) return set_err(GSS_E_STATE, "the DB is not tagged");
    if (e->pending || !e->inflight.empty())
        return set_err(GSS_E_STATE, "a sweep or step is in flight");
    HIP_TRY(hipStreamSynchronize(e->stream));
    std::vector<CompactRef> refs;
    std::vector<int> lens;
    int64_t N = 0;
    for (size_t len = 1; len < e->buckets.size(); len++) {
        const Bucket &b = e->buckets[len];
        if (!b.count) continue;
        refs.push_back({b.d, b.tags, nullptr, nullptr, N, (uint32_t)len, 0});
        lens.push_back((int)len);
        N += b.count;
    }
    if (!N) return GSS_OK;
    if (N > 0x7FFFFFFF) return set_err(GSS_E_INVAL, "DB too large to compact in one pass");
    const int nb = (int)refs.size();
    size_t temp_bytes = 0;
    HIP_TRY(scan_exclusive_u32(nullptr, nullptr, N, nullptr, temp_bytes, e->stream));
    if ((rc = e->d_cref.reserve(nb)) || (rc = e->d_flags.reserve(N)) ||
        (rc = e->d_pos.reserve(N)) || (rc = e->d_kept.reserve(nb)) ||
        (rc = e->h_kept.reserve(nb)) || (rc = e->d_scan_temp.reserve(std::max<size_t>(temp_bytes, 1))))
        return rc;
    HIP_TRY(hipMemcpyAsync(e->d_cref.p, refs.data(), nb * sizeof(CompactRef),
                           hipMemcpyHostToDevice, e->stream));
    HIP_TRY(launch_compact_flags(e->d_cref.p, nb, N, d_keep, ntags, e->d_flags.p, e->stream));
    HIP_TRY(scan_exclusive_u32(e->d_flags.p, e->d_pos.p, N, e->d_scan_temp.p, temp_bytes,
                               e->stream));
    HIP_TRY(launch_compact_bounds(e->d_cref.p, nb, N, e->d_flags.p, e->d_pos.p, e->d_kept.p,
                                  e->stream));
    HIP_TRY(hipMemcpyAsync(e->h_kept.p, e->d_kept.p, nb * sizeof(uint32_t),
                           hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    std::vector<int64_t> caps(nb);
    for (int i = 0; i < nb; i++) {
        const int64_t k = e->h_kept.p[i];
        caps[i] = std::max<int64_t>(k + k / 4, 256);
        uint32_t *dl = nullptr, *dt = nullptr;
        hipError_t he = hipMallocAsync((void **)&dl, (size_t)caps[i] * lens[i] * 4, e->stream);
        if (he == hipSuccess) he = hipMallocAsync((void **)&dt, (size_t)caps[i] * 4, e->stream);
        if (he != hipSuccess) {
            if (dl) (void)hipFreeAsync(dl, e->stream);
            for (int j = 0; j < i; j++) {
                (void)hipFreeAsync(refs[j].dst_lits, e->stream);
                (void)hipFreeAsync(refs[j].dst_tags, e->stream);
            }
            return set_err(GSS_E_NOMEM, "compaction: allocation failed");
        }
        refs[i].dst_lits = dl;
        refs[i].dst_tags = dt;
    }
    HIP_TRY(hipMemcpyAsync(e->d_cref.p, refs.data(), nb * sizeof(CompactRef),
                           hipMemcpyHostToDevice, e->stream));
    HIP_TRY(launch_compact_move(e->d_cref.p, nb, N, e->d_flags.p, e->d_pos.p, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    e->rows_version++;
    e->state_version++;
    for (int i = 0; i < nb; i++) {
        Bucket &b = e->buckets[lens[i]];
        (void)hipFreeAsync(b.d, e->stream);
        (void)hipFreeAsync(b.tags, e->stream);
        const int64_t k = e->h_kept.p[i];
        e->total_clauses -= b.count - k;
        e->total_lits -= (b.count - k) * lens[i];
        b.d = refs[i].dst_lits;
        b.tags = refs[i].dst_tags;
        b.cap = caps[i];
        b.count = k;
    }
    return GSS_OK;
}

int gss::db_tag_arrays(gss_engine *e, std::vector<const uint32_t *> &tags) {
    if (!e->tagged) return set_err(GSS_E_STATE, "the DB is not tagged");
    tags.assign(e->buckets.size(), nullptr);
    for (size_t len = 1; len < e->buckets.size(); len++) tags[len] = e->buckets[len].tags;
    return GSS_OK;
}

int gss::sweep_records(gss_engine *e, const gss_hit **hits, const uint32_t **fills,
                       uint32_t *chunks) {
    if (!e->waited) return set_err(GSS_E_STATE, "no waited sweep");
    if (e->sweep_rows_version != e->rows_version)
        return set_err(GSS_E_STATE, "rows moved since the last sweep");
    *hits = e->d_hits.p;
    *fills = e->d_fills.p;
    *chunks = e->last_chunks;
    return GSS_OK;
}

hipStream_t gss::engine_stream(gss_engine *e) { return e->stream; }

int gss::engine_activate(gss_engine *e) { return check_engine(e); }

extern "C" {

int gss_db_compact(gss_engine *e, const uint8_t *keep, int64_t ntags) {
    int rc = check_engine(e);
    if (rc) return rc;
    if (ntags < 0 || (ntags && !keep)) return set_err(GSS_E_INVAL, "bad keep flags");
    if ((rc = e->d_keep.reserve(std::max<int64_t>(ntags, 1)))) return rc;
    if (ntags)
        HIP_TRY(hipMemcpyAsync(e->d_keep.p, keep, (size_t)ntags, hipMemcpyHostToDevice,
                               e->stream));
    return gss::db_compact_device(e, e->d_keep.p, ntags);
}

int gss_db_set_bucket(gss_engine *e, int len, int64_t n, const uint32_t *lits) {
    int rc = check_engine(e);
    if (rc) return rc;
    if ((rc = settle_steps(e))) return rc;
    if (e->tagged) return set_err(GSS_E_STATE, "tagged DB: compact with gss_db_compact");
    if (len < 1 || len > GSS_MAX_CLAUSE_LEN || n < 0 || (n > 0 && !lits))
        return set_err(GSS_E_INVAL, "bad bucket replacement");
    uint32_t mx = 0;
    for (int64_t i = 0; i < n * len; i++) mx = std::max(mx, lits[i]);
    if (mx > kMaxLit) return set_err(GSS_E_INVAL, "literal %u exceeds %u", mx, kMaxLit);
    if ((int)e->buckets.size() <= len) e->buckets.resize(len + 1);
    Bucket &b = e->buckets[len];
    const int64_t old = b.count;
    e->rows_version++;
    e->state_version++;
    if (n == 0) {
        // emptying a bucket allocates nothing (out-of-memory recovery relies
        // on it): its array goes back to the pool in stream order
        if (b.d) HIP_TRY(hipFreeAsync(b.d, e->stream));
        b.d = nullptr;
        b.cap = 0;
        b.count = 0;
        e->total_clauses -= old;
        e->total_lits -= old * len;
        return GSS_OK;
    }
    b.count = 0;  // nothing to preserve when growing
    if ((rc = ensure_bucket_cap(e, len, n, true))) {
        b.count = old;
        return rc;
    }
    if (n) {
        std::vector<uint32_t> stored((size_t)n * len);
        for (size_t i = 0; i < stored.size(); i++) stored[i] = lits[i] << kLitShift;
        HIP_TRY(hipMemcpyAsync(b.d, stored.data(), stored.size() * sizeof(uint32_t),
                               hipMemcpyHostToDevice, e->stream));
        HIP_TRY(hipStreamSynchronize(e->stream));
    }
    b.count = n;
    e->total_clauses += n - old;
    e->total_lits += (n - old) * len;
    if (n) e->max_lit = std::max<int64_t>(e->max_lit, mx);
    return GSS_OK;
}

int gss_db_bucket_count(gss_engine *e, int len, int64_t *out) {
    if (!e || !out) return set_err(GSS_E_INVAL, "null argument");
    *out = (len >= 1 && len < (int)e->buckets.size()) ? e->buckets[len].count : 0;
    return GSS_OK;
}

int gss_db_totals(gss_engine *e, int64_t *clauses, int64_t *literals) {
    if (!e) return set_err(GSS_E_INVAL, "null engine");
    if (clauses) *clauses = e->total_clauses;
    if (literals) *literals = e->total_lits;
    return GSS_OK;
}

int gss_db_read(gss_engine *e, int len, int64_t row0, int64_t n, uint32_t *out) {
    int rc = check_engine(e);
    if (rc) return rc;
    if ((rc = settle_steps(e))) return rc;
    if (n == 0 && row0 == 0 && len >= 1) return GSS_OK;  // nothing to read
    if (len < 1 || len >= (int)e->buckets.size() || row0 < 0 || n < 0 ||
        row0 + n > e->buckets[len].count || (n && !out))
        return set_err(GSS_E_INVAL, "bad bucket read");
    if (n)
        HIP_TRY(hipMemcpyAsync(out, e->buckets[len].d + (size_t)row0 * len,
                               (size_t)n * len * sizeof(uint32_t), hipMemcpyDeviceToHost,
                               e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    for (int64_t i = 0; i < n * len; i++) out[i] >>= kLitShift;  // stored offsets -> literals
    return GSS_OK;
}

int gss_db_clear(gss_engine *e) {
    int rc = check_engine(e);
    if (rc) return rc;
    if ((rc = settle_steps(e))) return rc;
    for (Bucket &b : e->buckets) b.count = 0;
    e->rows_version++;
    e->state_version++;
    e->total_clauses = e->total_lits = 0;
    e->max_lit = -1;
    return GSS_OK;
}

/* ---- assignment table ----------------------------------------------------- */

// The table has 2 * (vars + 1) rows.  Row l holds the slots where literal l
// is NOT false (all ones = undefined everywhere): row 2v = v not false, row
// 2v+1 = v not true.  Two sentinel rows follow: row 2*vars is all zeros
// ("false in every slot") and row 2*vars + 1 all ones ("true in no slot");
// the sweep reads them for padding lanes past the end of a clause, so those
// lanes need no masking.
static size_t table_rows(int64_t vars) { return (size_t)(vars + 1) * 2; }

static int table_fits(int64_t vars, int words) {
    // the sweep addresses the table with 32-bit byte offsets
    if ((uint64_t)table_rows(vars) * (uint64_t)words * 8ull > 0xFFFFFFFFull)
        return set_err(GSS_E_INVAL, "assignment table of %lld vars x %d words exceeds 4 GiB",
                       (long long)vars, words);
    // every row, the two sentinels included, must have a 32-bit stored offset
    if (table_rows(vars) - 1 > (int64_t)kMaxLit)
        return set_err(GSS_E_INVAL, "%lld variables exceed the %u the table can address",
                       (long long)vars, (kMaxLit - 1) / 2);
    return GSS_OK;
}

static int write_sentinels(gss_engine *e) {
    HIP_TRY(launch_fill_u64(e->table.p + (size_t)e->vars * 2 * e->words, e->words, 0ull,
                            e->stream));
    HIP_TRY(launch_fill_u64(e->table.p + ((size_t)e->vars * 2 + 1) * e->words, e->words, ~0ull,
                            e->stream));
    return GSS_OK;
}

int gss_assig_reset(gss_engine *e, int64_t var_count, int words) {
    int rc = check_engine(e);
    if (rc) return rc;
    if ((rc = settle_steps(e))) return rc;
    if (var_count < 0 || var_count > (1ll << 31) - 2 || words < 1 || words > GSS_MAX_WORDS ||
        log2_exact(words) < 0)
        return set_err(GSS_E_INVAL, "bad table shape (%lld vars, %d words; words must be a power "
                                    "of two <= %d)",
                       (long long)var_count, words, GSS_MAX_WORDS);
    if ((rc = table_fits(var_count, words))) return rc;
    HIP_TRY(hipStreamSynchronize(e->stream));
    const size_t n = table_rows(var_count) * words;
    if (n > e->table.cap) {
        e->table.release();
        if ((rc = e->table.reserve(n))) return rc;
    }
    HIP_TRY(launch_fill_u64(e->table.p, (int64_t)n, ~0ull, e->stream));  // all undefined
    e->state_version++;
    e->vars = var_count;
    e->words = words;
    if ((rc = write_sentinels(e))) return rc;
    HIP_TRY(hipStreamSynchronize(e->stream));
    return GSS_OK;
}

int gss_assig_grow_vars(gss_engine *e, int64_t var_count) {
    int rc = check_engine(e);
    if (rc) return rc;
    if ((rc = settle_steps(e))) return rc;
    if (var_count < e->vars || var_count > (1ll << 31) - 2)
        return set_err(GSS_E_INVAL, "variable count can only grow");
    if ((rc = table_fits(var_count, e->words))) return rc;
    // rows of the old sentinels become the first new variable: clear from there
    const size_t old_n = (size_t)e->vars * 2 * e->words;
    const size_t n = table_rows(var_count) * e->words;
    if ((rc = e->table.reserve(n, true, e->stream))) return rc;
    HIP_TRY(launch_fill_u64(e->table.p + old_n, (int64_t)(n - old_n), ~0ull, e->stream));
    e->state_version++;
    e->vars = var_count;
    if ((rc = write_sentinels(e))) return rc;
    HIP_TRY(hipStreamSynchronize(e->stream));
    return GSS_OK;
}

int gss_assig_upload(gss_engine *e, const uint64_t *is_def, const uint64_t *is_true) {
    int rc = check_engine(e);
    if (rc) return rc;
    if ((rc = settle_steps(e))) return rc;
    if (!is_def || !is_true) return set_err(GSS_E_INVAL, "null plane");
    const int W = e->words;
    const int64_t V = e->vars;
    // stage per block of variables
    const int64_t blk = 1 << 16;
    e->state_version++;
    Pinned<uint64_t> st;
    if ((rc = st.reserve((size_t)blk * 2 * W))) return rc;
    for (int64_t v0 = 0; v0 < V; v0 += blk) {
        const int64_t nv = std::min(blk, V - v0);
        for (int64_t v = 0; v < nv; v++) {
            for (int w = 0; w < W; w++) {
                const uint64_t d = is_def[(v0 + v) * W + w], t = is_true[(v0 + v) * W + w];
                st.p[(v * 2) * W + w] = ~(d & ~t);
                st.p[(v * 2 + 1) * W + w] = ~(d & t);
            }
        }
        HIP_TRY(hipMemcpyAsync(e->table.p + (size_t)v0 * 2 * W, st.p,
                               (size_t)nv * 2 * W * sizeof(uint64_t), hipMemcpyHostToDevice,
                               e->stream));
        HIP_TRY(hipStreamSynchronize(e->stream));
    }
    return GSS_OK;
}

int gss_assig_download(gss_engine *e, uint64_t *is_def, uint64_t *is_true) {
    int rc = check_engine(e);
    if (rc) return rc;
    if ((rc = settle_steps(e))) return rc;
    if (!is_def || !is_true) return set_err(GSS_E_INVAL, "null plane");
    const int W = e->words;
    const int64_t V = e->vars;
    std::vector<uint64_t> t((size_t)V * 2 * W);
    if (V)
        HIP_TRY(hipMemcpyAsync(t.data(), e->table.p, t.size() * sizeof(uint64_t),
                               hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    for (int64_t v = 0; v < V; v++)
        for (int w = 0; w < W; w++) {
            const uint64_t f = ~t[(v * 2) * W + w], tr = ~t[(v * 2 + 1) * W + w];
            is_def[v * W + w] = f | tr;
            is_true[v * W + w] = tr;
        }
    return GSS_OK;
}

static int upload_updates(gss_engine *e, int64_t n, const gss_var_update *ups) {
    int rc;
    for (int64_t i = 0; i < n; i++)
        if (ups[i].var >= (uint64_t)e->vars || ups[i].word >= (uint32_t)e->words)
            return set_err(GSS_E_INVAL, "update %lld: var %u / word %u out of range",
                           (long long)i, ups[i].var, ups[i].word);
    HIP_TRY(hipStreamSynchronize(e->stream));
    e->applied = 0;
    if ((rc = e->h_ups.reserve((size_t)n)) || (rc = e->d_ups.reserve((size_t)n))) return rc;
    std::memcpy(e->h_ups.p, ups, (size_t)n * sizeof(gss_var_update));
    HIP_TRY(hipMemcpyAsync(e->d_ups.p, e->h_ups.p, (size_t)n * sizeof(gss_var_update),
                           hipMemcpyHostToDevice, e->stream));
    return GSS_OK;
}

int gss_assig_apply(gss_engine *e, int64_t n, const gss_var_update *ups) {
    int rc = check_engine(e);
    if (rc) return rc;
    if ((rc = settle_steps(e))) return rc;
    if (n < 0 || (n && !ups)) return set_err(GSS_E_INVAL, "bad update batch");
    if (!n) {
        e->applied = 0;
        return GSS_OK;
    }
    if ((rc = upload_updates(e, n, ups))) return rc;
    HIP_TRY(launch_assig_apply(e->d_ups.p, n, e->table.p, e->words, e->stream));
    e->applied = n;
    e->state_version++;
    return GSS_OK;
}

int gss_assig_roll_forward_applied(gss_engine *e, const int8_t *last_pos) {
    int rc = check_engine(e);
    if (rc) return rc;
    if ((rc = settle_steps(e))) return rc;
    if (!last_pos) return set_err(GSS_E_INVAL, "null last_pos");
    if (!e->applied) return GSS_OK;
    for (int w = 0; w < e->words; w++)
        if (last_pos[w] > 63) return set_err(GSS_E_INVAL, "last_pos[%d] = %d > 63", w, last_pos[w]);
    // h_lastpos is only rewritten after the stream has drained (sweep wait or
    // the synchronisation in upload_updates), so the async copy is safe
    std::memcpy(e->h_lastpos.p, last_pos, e->words);
    HIP_TRY(hipMemcpyAsync(e->d_lastpos, e->h_lastpos.p, e->words, hipMemcpyHostToDevice,
                           e->stream));
    HIP_TRY(launch_assig_roll(e->d_ups.p, e->applied, e->table.p, e->words, e->d_lastpos,
                              e->stream));
    e->applied = 0;
    e->state_version++;
    return GSS_OK;
}

int gss_assig_roll_forward(gss_engine *e, int64_t n, const gss_var_update *ups,
                           const int8_t *last_pos) {
    int rc = check_engine(e);
    if (rc) return rc;
    if ((rc = settle_steps(e))) return rc;
    if (n < 0 || (n && (!ups || !last_pos))) return set_err(GSS_E_INVAL, "bad roll-forward");
    if (!n) return GSS_OK;
    for (int w = 0; w < e->words; w++)
        if (last_pos[w] > 63) return set_err(GSS_E_INVAL, "last_pos[%d] = %d > 63", w, last_pos[w]);
    if ((rc = upload_updates(e, n, ups))) return rc;
    std::memcpy(e->h_lastpos.p, last_pos, e->words);
    HIP_TRY(hipMemcpyAsync(e->d_lastpos, e->h_lastpos.p, e->words, hipMemcpyHostToDevice,
                           e->stream));
    HIP_TRY(launch_assig_roll(e->d_ups.p, n, e->table.p, e->words, e->d_lastpos, e->stream));
    e->state_version++;
    HIP_TRY(hipStreamSynchronize(e->stream));
    return GSS_OK;
}

/* ---- sweep ---------------------------------------------------------------- */

int gss_sweep_async(gss_engine *e, const uint64_t *valid) {
    int rc = check_engine(e);
    if (rc) return rc;
    if ((rc = settle_steps(e))) return rc;
    if (!valid) return set_err(GSS_E_INVAL, "null valid mask");
    if (e->pending) return set_err(GSS_E_STATE, "previous sweep not waited for");
    if (e->total_clauses > 0 && e->max_lit >= 2 * e->vars)
        return set_err(GSS_E_STATE, "clause literal %lld needs variable %lld but the table has "
                                    "%lld variables",
                       (long long)e->max_lit, (long long)(e->max_lit >> 1), (long long)e->vars);
    const int W = e->words;
    const int cpp = sweep_clauses_per_pass(W);
    int tc = e->tune_task_clauses > 0 ? e->tune_task_clauses : 16 * cpp;
    tc = std::max(cpp, (tc / cpp) * cpp);
    e->task_clauses = tc;
    // descriptors of non-empty buckets
    int nb = 0;
    for (size_t len = 1; len < e->buckets.size(); len++)
        if (e->buckets[len].count) nb++;
    if ((rc = e->h_desc.reserve(std::max(nb, 1))) || (rc = e->d_desc.reserve(std::max(nb, 1))))
        return rc;
    int64_t tasks = 0;
    e->desc_new.clear();
    const size_t nlen = e->buckets.size();
    for (size_t i = 1; i < nlen; i++) {
        // tasks of long clauses come first (they are the most expensive), so
        // the persistent waves finish together; schedule 2 reverses that
        const size_t len = e->tune_schedule == 2 ? i : nlen - i;
        const Bucket &b = e->buckets[len];
        if (!b.count) continue;
        // the kernel does row arithmetic on the count as a signed 32-bit value
        if (b.count > 0x7FFFFFFFll)
            return set_err(GSS_E_INVAL, "bucket %zu holds %lld rows (at most 2^31 - 1)", len,
                           (long long)b.count);
        BucketDesc d;
        d.lits = b.d;
        d.len = (uint32_t)len;
        d.count = (uint32_t)b.count;
        d.task_begin = (uint32_t)tasks;
        d.pad = 0;
        e->desc_new.push_back(d);
        tasks += (b.count + tc - 1) / tc;
    }
    if (tasks > 0x7FFFFFFF) return set_err(GSS_E_INVAL, "too many tasks");
    e->nbuckets = nb;
    e->ntasks = (int)tasks;
    int active_words = 0;
    for (int w = 0; w < W; w++)
        if (valid[w]) active_words++;
    if (!active_words) e->ntasks = 0;  // every clause starts dead: nothing to check
    // Uploads only when what the device holds differs (the previous sweep has
    // been waited for, so its copies out of the pinned staging are done).
    const bool valid_same = (int)e->valid_up.size() == W &&
                            std::memcmp(e->valid_up.data(), valid, W * sizeof(uint64_t)) == 0;
    const bool desc_same =
        e->desc_up_at == e->d_desc.p && e->desc_up.size() == e->desc_new.size() &&
        (nb == 0 ||
         std::memcmp(e->desc_up.data(), e->desc_new.data(), nb * sizeof(BucketDesc)) == 0);
    if (!valid_same) {
        std::memcpy(e->h_valid.p, valid, W * sizeof(uint64_t));
        HIP_TRY(hipMemcpyAsync(e->d_valid, e->h_valid.p, W * sizeof(uint64_t),
                               hipMemcpyHostToDevice, e->stream));
        e->valid_up.assign(valid, valid + W);
    }
    if (!desc_same) {
        if (nb) {
            std::memcpy(e->h_desc.p, e->desc_new.data(), nb * sizeof(BucketDesc));
            HIP_TRY(hipMemcpyAsync(e->d_desc.p, e->h_desc.p, nb * sizeof(BucketDesc),
                                   hipMemcpyHostToDevice, e->stream));
        }
        e->desc_up = e->desc_new;
        e->desc_up_at = e->d_desc.p;
    }
    // persistent grid
    int per_cu = resident_blocks(e);
    if (e->tune_waves_per_cu > 0) per_cu = std::max(1, std::min(per_cu, e->tune_waves_per_cu / 4));
    int64_t blocks = (int64_t)e->cu * per_cu;
    const int64_t need_blocks = (tasks + 3) / 4;
    blocks = std::max<int64_t>(1, std::min(blocks, need_blocks));
    e->blocks = (int)blocks;
    // Hit capacity: an abandoned chunk holds more than kHitChunk - 64 records
    // and each wave leaves at most one partial chunk, so this bound is never
    // exceeded while records <= clauses * active words.
    const uint64_t ub = (uint64_t)e->total_clauses * (uint64_t)active_words;
    uint64_t chunks = ub / (kHitChunk - 64) + 1 + (uint64_t)blocks * 4;
    size_t free_b = 0, total_b = 0;
    if (chunks > e->chunk_cap && hipMemGetInfo(&free_b, &total_b) == hipSuccess) {
        // never take more than a quarter of the free memory; beyond that the
        // wait re-runs with exactly what the claims asked for
        const uint64_t budget =
            e->chunk_cap + (uint64_t)(free_b / 4) / (kHitChunk * sizeof(gss_hit));
        chunks = std::min(chunks, std::max<uint64_t>(budget, 1));
    }
    if (e->tune_max_chunks > 0) chunks = std::min<uint64_t>(chunks, e->tune_max_chunks);
    if (chunks > e->chunk_cap && (rc = reserve_chunks(e, chunks))) return rc;
    if ((rc = launch_current_sweep(e))) return rc;
    e->sweep_rows_version = e->rows_version;
    e->sweep_state_version = e->state_version;
    e->pending = true;
    e->waited = false;
    e->have_sweep = true;
    return GSS_OK;
}

int gss_sweep_wait(gss_engine *e, int64_t *n_hits) {
    int rc = check_engine(e);
    if (rc) return rc;
    if (!e->pending) return set_err(GSS_E_STATE, "no sweep in flight");
    for (;;) {
        HIP_TRY(hipStreamSynchronize(e->stream));
        const uint32_t used = e->h_count.p[0];
        if (used <= e->chunk_cap) {
            e->last_chunks = used;
            break;
        }
        // The buffer was too small: grow to what the claims asked for, re-run
        // -- on the same DB and table only.
        if (e->state_version != e->sweep_state_version) {
            e->pending = false;
            return set_err(GSS_E_NOMEM,
                           "the sweep's hit buffer overflowed (%u chunks claimed, %u held) and the "
                           "clause DB or assignment table changed after the sweep was queued, so "
                           "it cannot be re-run; wait for a sweep before changing them",
                           used, e->chunk_cap);
        }
        if ((rc = reserve_chunks(e, (uint64_t)used + used / 8 + 16))) {
            e->pending = false;
            return rc;
        }
        if ((rc = launch_current_sweep(e))) {
            e->pending = false;
            return rc;
        }
    }
    e->pending = false;
    e->waited = true;
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, e->ev0, e->ev1));
    e->last_ms = ms;
    e->total_ms += ms;
    e->nsweeps++;
    const int64_t tot = (int64_t)((uint64_t)e->h_count.p[1] | ((uint64_t)e->h_count.p[2] << 32));
    e->last_hits = tot;
    if (n_hits) *n_hits = tot;
    return GSS_OK;
}

int gss_sweep(gss_engine *e, const uint64_t *valid, int64_t *n_hits) {
    int rc = gss_sweep_async(e, valid);
    if (rc) return rc;
    return gss_sweep_wait(e, n_hits);
}

int gss_hits_copy(gss_engine *e, gss_hit *out, int64_t cap, int64_t *n) {
    int rc = check_engine(e);
    if (rc) return rc;
    if (!e->waited) return set_err(GSS_E_STATE, "no waited sweep");
    if (n) *n = e->last_hits;
    const int64_t want = std::min<int64_t>(cap, e->last_hits);
    if (want <= 0) return GSS_OK;
    if (!out) return set_err(GSS_E_INVAL, "null output");
    // Compact the claimed chunks into a dense device buffer (offsets and the
    // copy computed on the device), then one D2H.
    if ((rc = e->d_dense.reserve((size_t)e->last_hits)) ||
        (rc = e->d_offsets.reserve((size_t)e->chunk_cap + 1)))
        return rc;
    HIP_TRY(launch_compact_records(e->d_alloc, e->d_fills.p, e->chunk_cap, e->d_offsets.p,
                                   e->d_hits.p, e->d_dense.p, e->d_dense.cap, 3, e->stream));
    HIP_TRY(hipMemcpyAsync(out, e->d_dense.p, (size_t)want * sizeof(gss_hit),
                           hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    return GSS_OK;
}

int gss_hits_gather(gss_engine *e, gss_hit *d_out, int64_t cap, int64_t *n) {
    int rc = check_engine(e);
    if (rc) return rc;
    if (!e->waited) return set_err(GSS_E_STATE, "no waited sweep");
    if (n) *n = e->last_hits;
    if (!d_out && cap == 0) return GSS_OK;  // the count only
    if (!d_out || cap < 0) return set_err(GSS_E_INVAL, "null output");
    if (cap < e->last_hits)
        return set_err(GSS_E_INVAL, "the sweep has %lld records; the output holds %lld",
                       (long long)e->last_hits, (long long)cap);
    if (e->last_chunks > kGatherMaxChunks)
        return set_err(GSS_E_INVAL,
                       "the sweep claimed %u hit chunks; gss_hits_gather handles at most %u "
                       "(gss_hits_copy has no such limit)",
                       e->last_chunks, kGatherMaxChunks);
    if (!e->last_hits) return GSS_OK;
    // the same two kernels as gss_hits_copy: chunk offsets into the scratch,
    // records straight into the caller's array
    HIP_TRY(launch_compact_records(e->d_alloc, e->d_fills.p, e->chunk_cap, e->gather_scratch,
                                   e->d_hits.p, d_out, (uint64_t)cap, 3, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    return GSS_OK;
}

int gss_sweep_kernel_ms(gss_engine *e, float *ms) {
    if (!e || !ms) return set_err(GSS_E_INVAL, "null argument");
    *ms = e->last_ms;
    return GSS_OK;
}

int gss_sweep_kernel_stats(gss_engine *e, double *total_ms, int64_t *count, int reset) {
    if (!e) return set_err(GSS_E_INVAL, "null engine");
    if (total_ms) *total_ms = e->total_ms;
    if (count) *count = e->nsweeps;
    if (reset) {
        e->total_ms = 0;
        e->nsweeps = 0;
    }
    return GSS_OK;
}

int gss_set_tuning(gss_engine *e, int key, int value) {
    if (!e) return set_err(GSS_E_INVAL, "null engine");
    if (value < 0 && !(key == GSS_TUNE_FUSE_LEN && value == -1))
        return set_err(GSS_E_INVAL, "negative tuning value");
    switch (key) {
        case GSS_TUNE_WAVES_PER_CU: e->tune_waves_per_cu = value; return GSS_OK;
        case GSS_TUNE_TASK_CLAUSES: e->tune_task_clauses = value; return GSS_OK;
        case GSS_TUNE_UNROLL:
            if (value != 0 && value != 1 && value != 2 && value != 4 && value != 8)
                return set_err(GSS_E_INVAL, "unroll must be 0, 1, 2, 4 or 8");
            e->tune_unroll = value;
            return GSS_OK;
        case GSS_TUNE_FUSE_LEN: e->tune_fuse_len = value; return GSS_OK;
        case GSS_TUNE_SCHEDULE:
            if (value > 2) return set_err(GSS_E_INVAL, "schedule must be 0, 1 or 2");
            e->tune_schedule = value;
            return GSS_OK;
        case GSS_TUNE_MAX_HIT_CHUNKS:
            e->tune_max_chunks = value;
            if (value > 0) {
                // drop the current buffer so the next sweep starts this small
                e->d_hits.release();
                e->d_fills.release();
                e->chunk_cap = 0;
            }
            return GSS_OK;
        default: return set_err(GSS_E_INVAL, "unknown tuning key %d", key);
    }
}

int gss_autotune(gss_engine *e, const uint64_t *valid, int reps, int *best, float *best_ms) {
    int rc = check_engine(e);
    if (rc) return rc;
    if ((rc = settle_steps(e))) return rc;
    if (!valid || reps < 1) return set_err(GSS_E_INVAL, "bad autotune arguments");
    if (e->pending) return set_err(GSS_E_STATE, "a sweep is in flight");
    struct Cfg {
        int unroll, fuse, task, wpc, sched;
    };
    std::vector<Cfg> grid;
    // rows in flight per lane: a choice for 16 words and wider, fixed below
    std::vector<int> unrolls = e->words >= 16 ? std::vector<int>{2, 4} : std::vector<int>{0};
    const int fuses[] = {-1, 1, 4, 8};
    const int tasks[] = {0, 32, 128};
    const int wpcs[] = {0, 24};
    for (int u : unrolls)
        for (int f : fuses)
            for (int t : tasks)
                for (int w : wpcs)
                    for (int sc = 1; sc <= 2; sc++) grid.push_back({u, f, t, w, sc});
    const Cfg saved{e->tune_unroll, e->tune_fuse_len, e->tune_task_clauses, e->tune_waves_per_cu,
                    e->tune_schedule};
    auto apply = [&](const Cfg &c) {
        e->tune_unroll = c.unroll;
        e->tune_fuse_len = c.fuse;
        e->tune_task_clauses = c.task;
        e->tune_waves_per_cu = c.wpc;
        e->tune_schedule = c.sched;
    };
    double total0 = e->total_ms;
    int64_t count0 = e->nsweeps;
    Cfg best_cfg = saved;
    float best_t = 0.f;
    bool have = false;
    for (const Cfg &c : grid) {
        apply(c);
        std::vector<float> ts;
        for (int r = 0; r < reps + 1; r++) {
            if ((rc = gss_sweep(e, valid, nullptr))) {
                apply(saved);
                return rc;
            }
            if (r) ts.push_back(e->last_ms);  // first run warms up
        }
        std::sort(ts.begin(), ts.end());
        const float med = ts[ts.size() / 2];
        if (!have || med < best_t) {
            best_t = med;
            best_cfg = c;
            have = true;
        }
    }
    apply(best_cfg);
    // autotuning sweeps are not part of the caller's kernel statistics
    e->total_ms = total0;
    e->nsweeps = count0;
    if (best) {
        best[0] = best_cfg.unroll;
        best[1] = best_cfg.fuse;
        best[2] = best_cfg.task;
        best[3] = best_cfg.wpc;
        best[4] = best_cfg.sched;
    }
    if (best_ms) *best_ms = best_t;
    return GSS_OK;
}

}  // extern "C"

/* ---- batched steps -------------------------------------------------------- */

// Persistent grid of a step: the resident maximum of 4-wave blocks, whatever
// the tasks (waves without a task end at once), so a growing DB never
// re-captures the graph for its grid.
static int step_blocks(gss_engine *e, int mode) {
    int per_cu = resident_blocks(e, mode);
    if (e->tune_waves_per_cu > 0) per_cu = std::max(1, std::min(per_cu, e->tune_waves_per_cu / 4));
    return (int)std::max<int64_t>(1, (int64_t)e->cu * per_cu);
}

// Queues the copy of a waited step's `total` records (in `used` chunks) to the
// slot's pinned h_hits on the slot's stream: one DMA copy when the graph made
// them dense, else a compaction first (and the next capture's dense buffer is
// sized for this volume).
static int queue_records_copy(gss_engine *e, StepSlot &sl, int64_t total, uint32_t used) {
    int rc;
    const int rw = sl.res_words;  // 8-byte words per record
    if (!sl.copy) HIP_TRY(hipStreamCreateWithFlags(&sl.copy, hipStreamNonBlocking));
    if (!sl.ev_copied) HIP_TRY(hipEventCreateWithFlags(&sl.ev_copied, hipEventDisableTiming));
    hipStream_t cs = sl.copy;
    sl.h_at = (sl.h_at + 1) % sl.h_n;  // the buffer of the slot's h_n-th last wait
    const int j = sl.h_at;
    e->records_hw = std::max<uint64_t>(e->records_hw, (uint64_t)total);
    // the ring's buffers count 24-byte units: n records of this form take
    auto units = [rw](uint64_t n) { return (size_t)((n * (uint64_t)rw + 2) / 3); };
    const size_t need = units((uint64_t)total);
    if (2 * need > sl.h_hits().cap) {
        // grown four-fold once the records pass half of it, to at least 24
        // MB -- from the replacement allocated ahead when there is one
        // (a pinned allocation stalls its thread for milliseconds per 100
        // MB); the previous copy into this slot's pinned memory must end
        // before a buffer is replaced
        const auto g0 = std::chrono::steady_clock::now();
        if (sl.copy_queued) HIP_TRY(hipEventSynchronize(sl.ev_copied));
        sl.copy_queued = false;
        int st = sl.spare_state[j].load(std::memory_order_acquire);
        if (st == 1) {
            e->prefetch->wait(&sl.spare_state[j]);
            st = sl.spare_state[j].load(std::memory_order_acquire);
        }
        if (st == 2 && sl.h_spare[j].cap >= 2 * need) {
            sl.h_hits().swap_with(sl.h_spare[j]);
            e->prefetch->release(sl.h_spare[j]);
        } else {
            if (st == 2) e->prefetch->release(sl.h_spare[j]);
            if ((rc = sl.h_hits().reserve(std::max<size_t>(4 * need, (size_t)1 << 20))))
                return rc;
        }
        sl.spare_state[j].store(0, std::memory_order_release);
        const int64_t us = std::chrono::duration_cast<std::chrono::microseconds>(
                               std::chrono::steady_clock::now() - g0)
                               .count();
        gss::log_stall("record buffer grown in a wait", us / 1000.0);
        e->host_grows.fetch_add(1, std::memory_order_relaxed);
        e->host_grow_us.fetch_add(us, std::memory_order_relaxed);
        if (us > e->host_grow_us_max.load(std::memory_order_relaxed))
            e->host_grow_us_max.store(us, std::memory_order_relaxed);
    } else if (sl.spare_state[j].load(std::memory_order_acquire) == 2) {
        // a replacement allocated ahead is ready: it takes the buffer's place
        // now, before a step needs it, once the slot's copies have ended (the
        // last one, and so every earlier one on the slot's stream)
        if (sl.copy_queued) {
            HIP_TRY(hipEventSynchronize(sl.ev_copied));
            sl.copy_queued = false;
        }
        if (sl.h_spare[j].cap > sl.h_hits().cap) sl.h_hits().swap_with(sl.h_spare[j]);
        e->prefetch->release(sl.h_spare[j]);
        sl.spare_state[j].store(0, std::memory_order_release);
    }
    // every buffer of the slot that holds less than three times the most
    // records a step has made gets a replacement of six times that, allocated
    // ahead: ready long before a step's records pass half of the buffer
    const size_t hw = units(e->records_hw);
    for (int b = 0; b < sl.h_n; b++) {
        if (3 * hw <= sl.h_ring[b].cap || sl.spare_state[b].load(std::memory_order_acquire) != 0)
            continue;
        if (!e->prefetch) {
            try {
                e->prefetch.reset(new HostPrefetch(e->device));
            } catch (...) {
                e->prefetch.reset();  // no thread: the wait allocates as before
            }
        }
        if (!e->prefetch) break;
        e->prefetch->allocate(&sl.h_spare[b], &sl.spare_state[b],
                              std::max<size_t>(6 * hw, (size_t)1 << 20));
    }
    if ((uint64_t)total > sl.step_dense) {
        if ((rc = sl.d_dense.reserve_async((size_t)total, cs)) ||
            (rc = sl.h_cseg.reserve(used)) || (rc = sl.d_cseg.reserve(used)))
            return rc;
        int ns = 0;
        uint64_t acc = 0;
        for (uint32_t c = 0; c < used; c++) {
            const uint32_t f = sl.h_out.p[1 + c];
            if (!f) continue;
            ScatterSeg &sg = sl.h_cseg.p[ns++];
            sg.dst = (uint32_t *)((uint64_t *)sl.d_dense.p + acc * rw);
            sg.src = (uint64_t)c * kHitChunk * rw * 2;
            sg.count = (uint64_t)f * rw * 2;
            acc += f;
        }
        HIP_TRY(hipMemcpyAsync(sl.d_cseg.p, sl.h_cseg.p, ns * sizeof(ScatterSeg),
                               hipMemcpyHostToDevice, cs));
        HIP_TRY(launch_scatter(sl.d_cseg.p, ns, (const uint32_t *)sl.d_hits.p, 0, cs));
        sl.dense_want = std::max<uint64_t>(sl.dense_want, 2 * (uint64_t)total);
    }
    HIP_TRY(hipMemcpyAsync(sl.h_hits().p, sl.d_dense.p, (size_t)total * rw * sizeof(uint64_t),
                           hipMemcpyDeviceToHost, cs));
    HIP_TRY(hipEventRecord(sl.ev_copied, cs));
    sl.copy_queued = true;
    return GSS_OK;
}

// Chunks a step's hit buffer can hold (the smallest of its parts).
static uint32_t chunk_room(const StepSlot &sl, int rw) {
    return (uint32_t)std::min<size_t>(
        std::min<size_t>(sl.d_fills.cap, sl.d_hits.cap / ((size_t)kHitChunk * rw)),
        sl.h_out.cap ? sl.h_out.cap - 1 : 0);
}

// Publishes the chunks the step hit buffers hold (read by gss_step_stats
// from any thread).
static void note_hit_chunks(gss_engine *e) {
    int64_t c = 0;
    for (const StepSlot &sl : e->slots) c += (int64_t)std::min(sl.d_fills.cap, sl.h_out.cap);
    e->step_hit_chunks.store(c, std::memory_order_relaxed);
}

// The part of a step after its table updates, over the slot's buffers and
// the settings of its graph: the sweep, the roll forward (held back if the
// records overflowed), the chunk fills to the host and the records made
// dense on the device right after the sweep, so that reading them on the
// host is one DMA copy (no compute units taken from the next step's sweep;
// serving steps: grouped by assignment word, so each request's records are
// one range).
static hipError_t launch_step_tail(gss_engine *e, StepSlot &sl, hipStream_t s) {
    StepHeader *dh = sl.d_hdr.p;
    SweepArgs a{};
    a.buckets = sl.d_desc.p;
    a.dyn = &dh->nbuckets;  // {nbuckets, ntasks}
    a.task_clauses = sl.g_tc;
    a.table = e->table.p;
    a.valid = dh->valid;
    a.hits = (gss_hit *)sl.d_hits.p;
    a.fills = sl.d_fills.p;
    a.alloc = sl.d_alloc;
    a.chunk_cap = sl.chunk_cap;
    a.fuse_len = sl.g_fuse;
    a.sentinel = (uint32_t)(2 * e->vars) << kLitShift;
    a.rec_words = sl.rec_words;
    uint8_t *agg =
        (sl.g_serve && e->serve_prefilter) || sl.g_screen ? e->serve_agg.p : nullptr;
    a.agg = agg;
    a.agg_words = agg ? e->serve_agg_words.p : nullptr;
    a.agg_vote = e->serve_vote;
    hipError_t he = launch_reset_u32(sl.d_alloc, 1, s);
    if (he == hipSuccess) he = launch_sweep(a, e->words, sl.g_unroll, sl.g_blocks, s);
    if (he == hipSuccess)
        he = launch_assig_roll_dyn(sl.d_ups.p, &dh->nups, e->table.p, e->words, dh->last_pos,
                                   sl.d_alloc, sl.chunk_cap, e->d_halt, s);
    // screened steps keep the aggregate current through the roll forward
    if (he == hipSuccess && sl.g_screen)
        he = launch_table_agg_update(e->table.p, dh, sl.d_ups.p, nullptr, false, e->words, agg,
                                     nullptr, s);
    if (he == hipSuccess)
        he = launch_stage_out(sl.d_alloc, sl.d_fills.p, sl.chunk_cap, sl.h_out.p, s);
    if (he == hipSuccess && sl.d_dense.cap > 0)
        he = sl.g_serve ? launch_group_by_word(sl.d_alloc, sl.d_fills.p, sl.chunk_cap,
                                               sl.d_hits.p, sl.rec_words, sl.d_wcount.p,
                                               sl.d_dense.p, sl.d_dense.cap, sl.h_group.p, s)
                        : launch_compact_records(sl.d_alloc, sl.d_fills.p, sl.chunk_cap,
                                                 sl.d_offs.p, sl.d_hits.p, sl.d_dense.p,
                                                 sl.d_dense.cap, sl.rec_words, s);
    return he;
}

// Captures the step's fixed launch sequence over this slot's buffers into
// graph entry g (whose previous graph, if any, is destroyed first).
static int build_step_graph(gss_engine *e, StepSlot &sl, StepSlot::StepGraph &gr, int blocks,
                            int tc, int fuse, bool serve, bool screen, bool agg_full) {
    if (gr.exec) (void)hipGraphExecDestroy(gr.exec);
    if (gr.graph) (void)hipGraphDestroy(gr.graph);
    gr.exec = nullptr;
    gr.graph = nullptr;
    gr.key.clear();
    sl.g_blocks = blocks;
    sl.g_tc = tc;
    sl.g_fuse = fuse;
    sl.g_unroll = e->tune_unroll;
    sl.g_serve = serve;
    sl.g_screen = screen;
    sl.g_agg_full = agg_full;
    if (!e->capture_stream)
        HIP_TRY(hipStreamCreateWithFlags(&e->capture_stream, hipStreamNonBlocking));
    hipStream_t cs = e->capture_stream;
    HIP_TRY(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    StageArgs st{};
    st.h[kRegHeader] = (const uint32_t *)sl.h_hdr.p;
    st.h[kRegDesc] = (const uint32_t *)sl.h_desc.p;
    st.h[kRegClauses] = (const uint32_t *)sl.h_cls.p;
    st.h[kRegLits] = sl.h_lits.p;
    st.h[kRegUps] = (const uint32_t *)sl.h_ups.p;
    st.h[kRegPatches] = (const uint32_t *)sl.h_patches.p;
    st.h[kRegBaseUps] = (const uint32_t *)sl.h_bups.p;
    st.h[kRegLoads] = (const uint32_t *)sl.h_loads.p;
    st.d[kRegHeader] = (uint32_t *)sl.d_hdr.p;
    st.d[kRegDesc] = (uint32_t *)sl.d_desc.p;
    st.d[kRegClauses] = (uint32_t *)sl.d_cls.p;
    st.d[kRegLits] = sl.d_lits.p;
    st.d[kRegUps] = (uint32_t *)sl.d_ups.p;
    st.d[kRegPatches] = (uint32_t *)sl.d_patches.p;
    st.d[kRegBaseUps] = (uint32_t *)sl.d_bups.p;
    st.d[kRegLoads] = (uint32_t *)sl.d_loads.p;
    st.halt = e->d_halt;
    StepHeader *dh = sl.d_hdr.p;
    uint8_t *agg = (serve && e->serve_prefilter) || screen ? e->serve_agg.p : nullptr;
    hipError_t he = hipSuccess;
    if (he == hipSuccess) he = launch_stage_in(st, cs);
    if (he == hipSuccess)
        he = launch_scatter_clauses(sl.d_cls.p, &dh->nclauses, sl.d_lits.p, kLitShift, cs);
    if (he == hipSuccess && sl.g_bases)
        he = launch_base_apply_dyn(sl.d_bups.p, &dh->nbups, e->bases.p, e->bases_vw, cs);
    // serving: the published requests that fit are admitted here, on the
    // device, when the step starts (their loads, valid slots and counts)
    if (he == hipSuccess && serve)
        he = launch_serve_admit(e->serve, dh, sl.d_loads.p, sl.h_report.p, e->words, cs);
    if (he == hipSuccess && serve) he = launch_serve_stage_patches(e->serve, dh, cs);
    // a serving step's words follow its loads and the rest are invalid in it,
    // so it writes whole table lines
    if (he == hipSuccess && sl.g_bases)
        he = serve && e->words % 2 == 0
                 ? launch_words_load_lines(sl.d_loads.p, &dh->nloads, e->bases.p, e->bases_vw,
                                           e->table.p, e->words, e->vars, cs)
                 : launch_words_load_dyn(sl.d_loads.p, &dh->nloads, e->bases.p, e->bases_vw,
                                         e->table.p, e->words, e->vars, cs);
    if (he == hipSuccess && agg && serve)
        he = launch_serve_agg(e->serve, dh, sl.d_loads.p, e->bases.p, e->bases_vw, e->vars, agg,
                              e->serve_agg_words.p, cs);
    if (he == hipSuccess) he = launch_assig_apply_dyn(sl.d_ups.p, &dh->nups, e->table.p, e->words, cs);
    if (he == hipSuccess && sl.g_patches)
        he = serve ? launch_serve_patch(e->serve, dh, e->table.p, e->words, agg, cs)
                   : launch_assig_patch_dyn(sl.d_patches.p, &dh->npatches, e->table.p, e->words,
                                            cs);
    // batched steps: the aggregate of the table as the sweep will read it,
    // rebuilt, or refreshed where this step's updates and patches changed it
    if (he == hipSuccess && screen)
        he = agg_full ? launch_table_agg(e->table.p, dh, e->vars, e->words, agg,
                                         e->serve_agg_words.p, cs)
                      : launch_table_agg_update(e->table.p, dh, sl.d_ups.p, sl.d_patches.p,
                                                sl.g_patches, e->words, agg,
                                                e->serve_agg_words.p, cs);
    if (he == hipSuccess) he = launch_step_tail(e, sl, cs);
    hipGraph_t g = nullptr;
    hipError_t he2 = hipStreamEndCapture(cs, &g);
    if (he != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return set_err(GSS_E_HIP, "step capture failed: %s", hipGetErrorString(he));
    }
    if (he2 != hipSuccess) return set_err(GSS_E_HIP, "step capture: %s", hipGetErrorString(he2));
    gr.graph = g;
    HIP_TRY(hipGraphInstantiate(&gr.exec, gr.graph, nullptr, nullptr, 0));
    gr.blocks = blocks;
    gr.tc = tc;
    gr.fuse = fuse;
    gr.unroll = e->tune_unroll;
    gr.serve = serve;
    gr.screen = screen;
    gr.agg_full = agg_full;
    gr.dense = sl.d_dense.cap;
    e->graph_captures.fetch_add(1, std::memory_order_relaxed);
    return GSS_OK;
}

// Recovers step `at` of the steps in flight, whose sweep claimed `used`
// chunks, more than its buffer held (the steps before it kept their
// records).  Its graph left the table as the sweep read it and halted the
// engine, so the later step in flight (if any) staged nothing.  Grows the
// buffer to what the sweep claimed, re-runs the step's sweep and what follows
// it, and replays the later step.  If the buffer cannot grow, the table
// still rolls forward and the later step is replayed, but this step's
// records are lost (its wait says so).
static int recover_step(gss_engine *e, size_t at, uint32_t used) {
    StepSlot &sl = e->slots[e->inflight[at]];
    hipStream_t es = e->stream;
    HIP_TRY(hipStreamSynchronize(es));
    // a later step's submit may have moved bucket arrays (growth copies the
    // rows in stream order): the step's rows [0, count) are the same rows
    // there
    const int nb = sl.h_hdr.p->nbuckets;
    for (int i = 0; i < nb; i++) sl.h_desc.p[i].lits = e->buckets[sl.h_desc.p[i].len].d;
    if (nb)
        HIP_TRY(hipMemcpyAsync(sl.d_desc.p, sl.h_desc.p, (size_t)nb * sizeof(BucketDesc),
                               hipMemcpyHostToDevice, es));
    const int rw = sl.rec_words;
    const uint64_t want = (uint64_t)used + used / 8 + 16;
    int rc = GSS_OK;
    if (want > chunk_room(sl, rw) &&
        ((rc = sl.d_hits.reserve_async((size_t)want * kHitChunk * rw, es)) ||
         (rc = sl.d_fills.reserve_async((size_t)want, es)) || (rc = sl.h_out.reserve(want + 1))))
        sl.records_lost = true;
    if (!sl.records_lost) {
        sl.chunk_cap = chunk_room(sl, rw);
        if ((rc = sl.d_offs.reserve_async((size_t)sl.chunk_cap + 1, es))) sl.records_lost = true;
    }
    note_hit_chunks(e);
    HIP_TRY(launch_reset_u32(e->d_halt, 1, es));
    if (sl.records_lost) {
        // no room: roll the table forward (the gate passes with no limit)
        HIP_TRY(launch_assig_roll_dyn(sl.d_ups.p, &sl.d_hdr.p->nups, e->table.p, e->words,
                                      sl.d_hdr.p->last_pos, sl.d_alloc, UINT32_MAX, e->d_halt,
                                      es));
    } else {
        HIP_TRY(launch_step_tail(e, sl, es));
    }
    HIP_TRY(hipEventRecord(sl.t1, es));
    // the later step, which staged nothing, runs again
    for (size_t i = at + 1; i < e->inflight.size(); i++) {
        StepSlot &nx = e->slots[e->inflight[i]];
        HIP_TRY(hipEventRecord(nx.t0, es));
        HIP_TRY(hipGraphLaunch(nx.graphs[nx.cur].exec, es));
        HIP_TRY(hipEventRecord(nx.t1, es));
        nx.settled = false;
    }
    e->step_hit_want = std::max<uint64_t>(e->step_hit_want, 2 * (uint64_t)used);
    e->step_recoveries++;
    return sl.records_lost ? set_err(GSS_E_NOMEM, "a step's hit buffer overflowed (%u chunks) and "
                                                  "could not grow: its records are lost",
                                     used)
                           : GSS_OK;
}

// Plain calls that read or change the clause DB or the assignment table
// while steps are in flight first make sure that no step in flight has
// halted the engine (else they would run between that step and its
// recovery): a step whose buffer could overflow is waited for on the device
// and recovered if it did.  Free when every buffer holds the worst case.
static int settle_steps(gss_engine *e) {
    for (size_t i = 0; i < e->inflight.size(); i++) {
        StepSlot &sl = e->slots[e->inflight[i]];
        if (sl.settled) continue;
        if (sl.chunk_cap < sl.worst_chunks && !sl.records_lost) {
            HIP_TRY(hipEventSynchronize(sl.t1));
            const uint32_t used = sl.h_out.p[0];
            if (used > sl.chunk_cap) {
                int rc = recover_step(e, i, used);
                if (rc && !sl.records_lost) return rc;
                HIP_TRY(hipEventSynchronize(sl.t1));
            }
        }
        sl.settled = true;
    }
    return GSS_OK;
}

extern "C" {

int gss_step_enqueue(gss_engine *e, int64_t n, const uint32_t *lits, const int32_t *lens,
                     uint32_t *out_rows, int64_t n_ups, const gss_var_update *ups,
                     const uint64_t *valid, const int8_t *last_pos) {
    gss_step_desc d{};
    d.n_clauses = n;
    d.lits = lits;
    d.lens = lens;
    d.out_rows = out_rows;
    d.n_ups = n_ups;
    d.ups = ups;
    d.valid = valid;
    d.last_pos = last_pos;
    return gss_step_submit(e, &d);
}

int gss_bases_reset(gss_engine *e, int nbases) {
    int rc = check_engine(e);
    if (rc) return rc;
    if (nbases < 0 || nbases > 65535) return set_err(GSS_E_INVAL, "base count %d", nbases);
    if (!e->inflight.empty() || e->pending) return set_err(GSS_E_STATE, "work in flight");
    const int64_t vw = (e->vars + 63) / 64;
    const size_t n = (size_t)nbases * 2 * (size_t)std::max<int64_t>(vw, 1);
    if (nbases && (rc = e->bases.reserve(n))) return rc;
    if (nbases) HIP_TRY(launch_fill_u64(e->bases.p, (int64_t)n, 0ull, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    e->nbases = nbases;
    e->bases_vw = vw;
    return GSS_OK;
}

int gss_step_submit(gss_engine *e, const gss_step_desc *desc) {
    return gss::step_submit_ex(e, desc, false);
}

int gss_step_stats(gss_engine *e, gss_step_counters *out) {
    if (!e || !out) return set_err(GSS_E_INVAL, "null argument");
    out->reruns = e->step_recoveries.load(std::memory_order_relaxed);
    out->hit_chunks = e->step_hit_chunks.load(std::memory_order_relaxed);
    out->host_grows = e->host_grows.load(std::memory_order_relaxed);
    out->host_grow_ms = 1e-3 * (double)e->host_grow_us.load(std::memory_order_relaxed);
    out->host_grow_ms_max = 1e-3 * (double)e->host_grow_us_max.load(std::memory_order_relaxed);
    return GSS_OK;
}

}  // extern "C"

int gss::engine_set_record_buffers(gss_engine *e, int n) {
    if (!e || n < 1 || n > kMaxHostRecordBuffers)
        return set_err(GSS_E_INVAL, "record buffers per slot: 1 .. %d", kMaxHostRecordBuffers);
    if (!e->inflight.empty()) return set_err(GSS_E_STATE, "steps in flight");
    for (StepSlot &sl : e->slots) {
        sl.h_n = n;
        sl.h_at = std::min(sl.h_at, n - 1);
    }
    return GSS_OK;
}

int64_t gss::engine_graph_captures(const gss_engine *e) {
    return e ? e->graph_captures.load(std::memory_order_relaxed) : 0;
}

int gss::engine_serve_attach(gss_engine *e, const ServeReq *reqs, uint32_t req_cap,
                             const gss_var_patch *patches, uint64_t patch_cap,
                             const uint64_t *head) {
    int rc = check_engine(e);
    if (rc) return rc;
    if (!reqs || !patches || !head || !req_cap || !patch_cap)
        return set_err(GSS_E_INVAL, "bad serving ring");
    if (!e->inflight.empty()) return set_err(GSS_E_STATE, "steps in flight");
    if ((rc = e->serve_tail.reserve(1)) || (rc = e->serve_adm.reserve(kServeMaxRequests)) ||
        (rc = e->serve_cnt.reserve(1)) || (rc = e->serve_dpatches.reserve(patch_cap)))
        return rc;
    HIP_TRY(launch_fill_u64(e->serve_tail.p, 1, 0ull, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    ServeRing &r = e->serve;
    r.reqs = reqs;
    r.patches = patches;
    r.head = head;
    r.tail = e->serve_tail.p;
    r.adm = e->serve_adm.p;
    r.adm_count = e->serve_cnt.p;
    r.dpatches = e->serve_dpatches.p;
    r.req_cap = req_cap;
    r.patch_cap = patch_cap;
    e->serve_on = true;
    return GSS_OK;
}

int gss::step_word_groups(gss_engine *e, uint64_t *offs, bool *grouped) {
    if (!e || !offs || !grouped) return set_err(GSS_E_INVAL, "null argument");
    if (e->last_step_slot < 0) return set_err(GSS_E_STATE, "no waited step");
    const StepSlot &sl = e->slots[e->last_step_slot];
    // grouped only if the graph's dense buffer held them (else the records
    // reached the host in chunk order)
    *grouped = sl.serve && sl.h_group.p && sl.rec_total > 0 &&
               (uint64_t)sl.rec_total <= sl.step_dense && sl.h_group.p[GSS_MAX_WORDS + 1] == 1;
    if (*grouped) std::memcpy(offs, sl.h_group.p, (GSS_MAX_WORDS + 1) * sizeof(uint64_t));
    return GSS_OK;
}

int gss::step_serve_report(gss_engine *e, ServeReport *out) {
    if (!e || !out) return set_err(GSS_E_INVAL, "null argument");
    if (e->last_step_slot < 0) return set_err(GSS_E_STATE, "no waited step");
    const StepSlot &sl = e->slots[e->last_step_slot];
    if (!sl.serve || !sl.h_report.p) return set_err(GSS_E_STATE, "not a serving step");
    *out = *sl.h_report.p;
    return GSS_OK;
}

int gss::step_submit_ex(gss_engine *e, const gss_step_desc *desc, bool serve) {
    int rc = check_engine(e);
    if (rc) return rc;
    if (!desc) return set_err(GSS_E_INVAL, "null step");
    if (serve && !e->serve_on) return set_err(GSS_E_STATE, "no serving ring attached");
    if (serve && (desc->n_loads || desc->n_patches || desc->n_ups))
        return set_err(GSS_E_INVAL, "a serving step takes its loads and patches from the ring "
                                    "and no updates");
    const int64_t n = desc->n_clauses, n_ups = desc->n_ups, n_patches = desc->n_patches,
                  n_bups = desc->n_base_ups, n_loads = desc->n_loads;
    const uint32_t *lits = desc->lits;
    const int32_t *lens = desc->lens;
    uint32_t *out_rows = desc->out_rows;
    const gss_var_update *ups = desc->ups;
    const uint64_t *valid = desc->valid;
    const int8_t *last_pos = desc->last_pos;
    if (n < 0 || (n > 0 && (!lits || !lens)) || n_ups < 0 || (n_ups > 0 && !ups) || !valid ||
        n_patches < 0 || (n_patches > 0 && !desc->patches) || n_bups < 0 ||
        (n_bups > 0 && !desc->base_ups) || n_loads < 0 || (n_loads > 0 && !desc->loads))
        return set_err(GSS_E_INVAL, "bad step arguments");
    if (desc->records != GSS_RECORDS_FULL && desc->records != GSS_RECORDS_IDS &&
        desc->records != GSS_RECORDS_CLAUSES)
        return set_err(GSS_E_INVAL, "records must be GSS_RECORDS_FULL, _IDS or _CLAUSES");
    if (serve && desc->records == GSS_RECORDS_CLAUSES)
        return set_err(GSS_E_INVAL, "serving steps group records by word: full or id records");
    if (desc->flags & ~GSS_STEP_SCREEN) return set_err(GSS_E_INVAL, "unknown step flags");
    // serving steps screen against their requests' bases instead
    const bool screen = !serve && (desc->flags & GSS_STEP_SCREEN);
    if (e->pending) return set_err(GSS_E_STATE, "a plain sweep is in flight");
    if (e->tagged && n > 0) return set_err(GSS_E_STATE, "tagged DB: append with gss_db_append_tagged");
    const int S = e->next_slot;
    StepSlot &sl = e->slots[S];
    if (sl.busy) return set_err(GSS_E_STATE, "two steps in flight: wait for the oldest first");
    const int W = e->words;
    if (n > 0x7FFFFFFF) return set_err(GSS_E_INVAL, "too many clauses in one step");
    // ---- validation (nothing changes before it passes).  The updates and
    // the literals are copied into this idle slot's pinned staging while they
    // are checked; a large step splits those copies over the stage pool.
    int maxlen = 0;
    int64_t total = 0;
    for (int64_t i = 0; i < n; i++) {
        if (lens[i] < 1 || lens[i] > GSS_MAX_CLAUSE_LEN)
            return set_err(GSS_E_INVAL, "clause %lld: length %d outside [1, %d]", (long long)i,
                           lens[i], GSS_MAX_CLAUSE_LEN);
        maxlen = std::max(maxlen, (int)lens[i]);
        total += lens[i];
    }
    if (total > 0xFFFFFFFFll) return set_err(GSS_E_INVAL, "too many literals in one step");
    if ((rc = sl.h_ups.reserve(std::max<int64_t>(n_ups, 1))) ||
        (rc = sl.h_lits.reserve(std::max<int64_t>(total, 1))))
        return rc;
    gss::StagePool *pool = nullptr;
    const size_t stage_bytes = (size_t)n_ups * sizeof(gss_var_update) + (size_t)total * 4 +
                               (size_t)n * sizeof(ClauseSeg);
    if (e->stage_threads > 1 && stage_bytes >= kStagePoolMinBytes) {
        if (!e->stage_pool) {
            try {
                e->stage_pool.reset(new gss::StagePool(e->stage_threads));
            } catch (...) {
                e->stage_threads = 0;  // no threads to be had: stage on this thread
            }
        }
        pool = e->stage_pool.get();
    }
    const int parts = pool ? pool->size() : 1;
    const uint32_t vmax = (uint32_t)std::min<int64_t>(e->vars, 0xFFFFFFFFll);
    std::vector<uint32_t> part_bad(parts, 0), part_max(parts, 0);
    // part k: updates [n_ups k / parts, ...) and literals [total k / parts, ...)
    auto copy_part = [&](int k) {
        const int64_t u0 = n_ups * k / parts, u1 = n_ups * (k + 1) / parts;
        uint32_t bad = 0;
        for (int64_t i = u0; i < u1; i++) {
            const gss_var_update u = ups[i];
            bad |= (uint32_t)(u.var >= vmax) | (uint32_t)(u.word >= (uint32_t)W);
            sl.h_ups.p[i] = u;
        }
        const int64_t l0 = total * k / parts, l1 = total * (k + 1) / parts;
        if (l1 > l0) std::memcpy(sl.h_lits.p + l0, lits + l0, (size_t)(l1 - l0) * sizeof(uint32_t));
        uint32_t mx = 0;
        for (int64_t i = l0; i < l1; i++) mx = std::max(mx, lits[i]);
        part_bad[k] = bad;
        part_max[k] = mx;
    };
    gss::StageJoin join;  // every return below waits for the copies first
    if (pool) {
        pool->start(copy_part);
        join.pool = pool;
    } else {
        copy_part(0);
    }
    for (int64_t i = 0; i < n_patches; i++)
        if (desc->patches[i].var >= (uint64_t)e->vars || desc->patches[i].word >= (uint32_t)W)
            return set_err(GSS_E_INVAL, "patch %lld out of range", (long long)i);
    if ((n_bups || n_loads || serve) && e->bases_vw != (e->vars + 63) / 64)
        return set_err(GSS_E_STATE, "the table changed size: reset the bases");
    for (int64_t i = 0; i < n_bups; i++)
        if (desc->base_ups[i].var >= (uint64_t)e->vars || desc->base_ups[i].base >= e->nbases)
            return set_err(GSS_E_INVAL, "base update %lld out of range", (long long)i);
    uint64_t loaded = 0;
    for (int64_t i = 0; i < n_loads; i++) {
        const gss_word_load &ld = desc->loads[i];
        if (ld.word >= W || ld.base >= e->nbases)
            return set_err(GSS_E_INVAL, "word load %lld out of range", (long long)i);
        if ((loaded >> ld.word) & 1) return set_err(GSS_E_INVAL, "word %u loaded twice", ld.word);
        loaded |= 1ull << ld.word;
    }
    int active_words = 0;
    for (int w = 0; w < W; w++) {
        if (last_pos && last_pos[w] > 63) return set_err(GSS_E_INVAL, "last_pos > 63");
        if (valid[w]) active_words++;
    }
    if (serve) active_words = W;  // any word may be admitted on the device
    // clause records: part k of cparts writes clauses [n k / cparts, ...); its
    // first row in each bucket (after the step's earlier parts) and its first
    // literal are counted here, while the copies run
    const int cparts = (pool && maxlen <= 4096 && n >= 4096) ? parts : 1;
    const size_t stride = (size_t)maxlen + 1;
    std::vector<int64_t> per(stride, 0), part_row(cparts * stride, 0);
    std::vector<uint32_t> part_pos(cparts, 0);
    {
        uint32_t pos = 0;
        for (int k = 0; k < cparts; k++) {
            int64_t *h = part_row.data() + k * stride;
            part_pos[k] = pos;
            const int64_t c0 = n * k / cparts, c1 = n * (k + 1) / cparts;
            for (int64_t i = c0; i < c1; i++) {
                h[lens[i]]++;
                pos += (uint32_t)lens[i];
            }
        }
        for (size_t len = 1; len < stride; len++) {
            int64_t acc = 0;
            for (int k = 0; k < cparts; k++) {
                int64_t &h = part_row[k * stride + len];
                const int64_t c = h;
                h = acc;
                acc += c;
            }
            per[len] = acc;
        }
    }
    if (pool) pool->wait();
    uint32_t bad = 0, mx = 0;
    for (int k = 0; k < parts; k++) {
        bad |= part_bad[k];
        mx = std::max(mx, part_max[k]);
    }
    if (bad)
        for (int64_t i = 0; i < n_ups; i++)
            if (ups[i].var >= (uint64_t)e->vars || ups[i].word >= (uint32_t)W)
                return set_err(GSS_E_INVAL, "update %lld: var %u / word %u out of range",
                               (long long)i, ups[i].var, ups[i].word);
    if (mx > kMaxLit) return set_err(GSS_E_INVAL, "literal %u exceeds %u", mx, kMaxLit);
    const int64_t need_lit = std::max<int64_t>(n ? (int64_t)mx : -1, e->max_lit);
    if (need_lit >= 2 * e->vars)
        return set_err(GSS_E_STATE, "clause literal %lld needs variable %lld but the table has "
                                    "%lld variables; grow the table before the step",
                       (long long)need_lit, (long long)(need_lit >> 1), (long long)e->vars);
    // ---- capacity: buckets, staging, device copies, hit buffer
    if ((int)e->buckets.size() <= maxlen) e->buckets.resize(maxlen + 1);  // empty buckets
    bool grow = false;
    for (int len = 1; len <= maxlen; len++)
        if (per[len] && e->buckets[len].count + per[len] > e->buckets[len].cap) grow = true;
    if (grow) {
        // bucket arrays move in stream order (ensure_bucket_cap): steps in
        // flight finish with the old arrays before they are released
        for (int len = 1; len <= maxlen; len++)
            if (per[len] && (rc = ensure_bucket_cap(e, len, e->buckets[len].count + per[len])))
                return rc;
    }
    const int cpp = sweep_clauses_per_pass(W);
    int tc = e->tune_task_clauses > 0 ? e->tune_task_clauses : 16 * cpp;
    tc = std::max(cpp, (tc / cpp) * cpp);
    const size_t nlen = std::max(e->buckets.size(), per.size());
    auto rows_after = [&](size_t len) -> int64_t {
        const int64_t c = len < e->buckets.size() ? e->buckets[len].count : 0;
        return c + (len < per.size() ? per[len] : 0);
    };
    int nb = 0;
    int64_t tasks = 0;
    for (size_t len = 1; len < nlen; len++) {
        const int64_t c = rows_after(len);
        if (c > 0x7FFFFFFFll)
            return set_err(GSS_E_INVAL, "bucket %zu would hold %lld rows (at most 2^31 - 1)", len,
                           (long long)c);
        if (c) nb++;
        tasks += (c + tc - 1) / tc;
    }
    if (tasks > 0x7FFFFFFF) return set_err(GSS_E_INVAL, "too many tasks");
    const int blocks = step_blocks(e, ((serve && e->serve_prefilter) || screen ? 1 : 0) +
                                          (desc->records == GSS_RECORDS_CLAUSES ? 2 : 0));
    // an abandoned chunk holds more than kHitChunk - 64 records and each wave
    // leaves at most one partial chunk: records <= clauses * active words
    // (one record per clause in the clause form) fit in `worst` chunks
    const uint64_t ub = (uint64_t)(e->total_clauses + n) *
                        (uint64_t)(desc->records == GSS_RECORDS_CLAUSES ? std::min(active_words, 1)
                                                                         : active_words);
    const uint64_t worst = ub / (kHitChunk - 64) + 1 + (uint64_t)blocks * 4;
    if ((rc = sl.h_hdr.reserve(1)) || (rc = sl.h_cls.reserve(std::max<int64_t>(n, 1))) ||
        (rc = sl.h_desc.reserve(std::max(nb, 1))) ||
        (rc = sl.h_patches.reserve(std::max<int64_t>(n_patches, 1))) ||
        (rc = sl.h_bups.reserve(std::max<int64_t>(n_bups, 1))) ||
        (rc = sl.h_loads.reserve(std::max<int64_t>(n_loads, 1))))
        return rc;
    // the slot's device copies are read only by its own graph, on the engine
    // stream: they grow in stream order, without stalling the other slot.
    // The copy of this slot's previous records (its own stream) still reads
    // d_hits and d_dense: the engine stream waits for it before anything of
    // this step, frees of replaced buffers included, runs there.
    hipStream_t es = e->stream;
    if (sl.copy_queued) HIP_TRY(hipStreamWaitEvent(es, sl.ev_copied, 0));
    if (serve && ((rc = sl.h_report.reserve(1)) || (rc = sl.h_group.reserve(GSS_MAX_WORDS + 2))))
        return rc;
    if (serve && (rc = sl.d_wcount.reserve_async(2 * GSS_MAX_WORDS, e->stream))) return rc;
    const bool use_agg = (serve && e->serve_prefilter) || screen;
    // a screened step rebuilds the aggregate unless the last change to the
    // table was a screened step's (word loads change whole words: rebuilt)
    const bool agg_full =
        screen && (e->table_agg_at == 0 || e->table_agg_at != e->state_version || n_loads > 0);
    if (use_agg && ((rc = e->serve_agg.reserve_async((size_t)((2 * e->vars + 2 + 3) / 4 * 4), es)) ||
                    (rc = e->serve_agg_words.reserve_async(8, es))))
        return rc;
    if ((rc = sl.d_hdr.reserve_async(1, es)) || (rc = sl.d_lits.reserve_async(sl.h_lits.cap, es)) ||
        (rc = sl.d_cls.reserve_async(sl.h_cls.cap, es)) ||
        (rc = sl.d_ups.reserve_async(sl.h_ups.cap, es)) ||
        (rc = sl.d_desc.reserve_async(sl.h_desc.cap, es)) ||
        (rc = sl.d_patches.reserve_async(sl.h_patches.cap, es)) ||
        (rc = sl.d_bups.reserve_async(sl.h_bups.cap, es)) ||
        (rc = sl.d_loads.reserve_async(std::max<size_t>(sl.h_loads.cap, GSS_MAX_WORDS), es)))
        return rc;
    // The hit buffer holds chunks of records of this step's form, as many as
    // twice the most a step has claimed (at least a floor: a partial chunk
    // per wave and 4M records, 16M for a serving engine), never more than the
    // worst case; it grows to twice that, so that a growing DB regrows it
    // rarely.  A step whose records overflow it is recovered (settle_steps,
    // recover_step).
    const int rw = desc->records == GSS_RECORDS_IDS ? 1 : desc->records == GSS_RECORDS_CLAUSES ? 2 : 3;
    const uint64_t floor_chunks = (serve ? 16384 : 4096) + (uint64_t)blocks * 4;
    uint64_t want = std::min(worst, std::max(floor_chunks, e->step_hit_want));
    if (e->tune_max_chunks > 0) want = std::min<uint64_t>(want, e->tune_max_chunks);
    if (want > chunk_room(sl, rw)) {
        const uint64_t chunks = e->tune_max_chunks > 0 ? want : 2 * want;
        if ((rc = sl.d_hits.reserve_async((size_t)chunks * kHitChunk * rw, e->stream)) ||
            (rc = sl.d_fills.reserve_async((size_t)chunks, e->stream)) ||
            (rc = sl.h_out.reserve(chunks + 1)))
            return rc;
        note_hit_chunks(e);
    }
    // (GSS_TUNE_MAX_HIT_CHUNKS holds every step to that many chunks)
    sl.chunk_cap = e->tune_max_chunks > 0
                       ? (uint32_t)std::min<uint64_t>(chunk_room(sl, rw), e->tune_max_chunks)
                       : chunk_room(sl, rw);
    sl.worst_chunks = worst;
    sl.rec_words = rw;
    if ((rc = sl.d_offs.reserve_async((size_t)sl.chunk_cap + 1, es))) return rc;
    if (sl.dense_want > sl.d_dense.cap && (rc = sl.d_dense.reserve_async(sl.dense_want, es)))
        return rc;
    if (!sl.d_alloc) HIP_TRY(hipMalloc((void **)&sl.d_alloc, 64));
    if (!sl.t0) HIP_TRY(hipEventCreate(&sl.t0));
    if (!sl.t1) HIP_TRY(hipEventCreate(&sl.t1));
    // ---- graph over exactly these buffers and settings
    const int fuse = fuse_of(e);
    sl.g_bases |= serve || n_bups > 0 || n_loads > 0;
    sl.g_patches |= serve || n_patches > 0;
    std::vector<const void *> key = {
        sl.h_hdr.p, sl.h_desc.p, sl.h_cls.p, sl.h_lits.p, sl.h_ups.p, sl.h_out.p,
        sl.h_patches.p, sl.h_bups.p, sl.h_loads.p,
        sl.d_hdr.p, sl.d_desc.p, sl.d_cls.p, sl.d_lits.p, sl.d_ups.p, sl.d_hits.p,
        sl.d_patches.p, sl.d_bups.p, sl.d_loads.p,
        sl.d_fills.p, sl.d_offs.p, sl.d_dense.p, (const void *)(intptr_t)sl.d_dense.cap,
        e->table.p, e->bases.p, (const void *)(intptr_t)e->bases_vw,
        (const void *)(intptr_t)W, (const void *)(intptr_t)e->vars,
        (const void *)(intptr_t)e->tune_unroll, (const void *)(intptr_t)blocks,
        (const void *)(intptr_t)tc, (const void *)(intptr_t)fuse,
        (const void *)(intptr_t)sl.chunk_cap, (const void *)(intptr_t)sl.rec_words,
        (const void *)(intptr_t)serve, sl.h_report.p, serve ? (const void *)e->serve.reqs : nullptr,
        sl.h_group.p, sl.d_wcount.p, (const void *)(intptr_t)(sl.g_bases + 2 * sl.g_patches),
        use_agg ? (const void *)e->serve_agg.p : nullptr,
        use_agg ? (const void *)e->serve_agg_words.p : nullptr,
        (const void *)(intptr_t)(screen ? (agg_full ? 1 : 2) : 0)};
    int gi = -1;
    for (size_t i = 0; i < sl.graphs.size(); i++)
        if (sl.graphs[i].exec && sl.graphs[i].key == key) gi = (int)i;
    if (gi < 0) {
        if (sl.graphs.size() < StepSlot::kGraphsPerSlot) {
            sl.graphs.emplace_back();
            gi = (int)sl.graphs.size() - 1;
        } else {
            gi = 0;
            for (size_t i = 1; i < sl.graphs.size(); i++)
                if (sl.graphs[i].used < sl.graphs[gi].used) gi = (int)i;
        }
        StepSlot::StepGraph &gr = sl.graphs[gi];
        const auto t0 = std::chrono::steady_clock::now();
        if ((rc = build_step_graph(e, sl, gr, blocks, tc, fuse, serve, screen, agg_full))) return rc;
        gss::log_stall("step graph capture", std::chrono::duration<double, std::milli>(
                                                  std::chrono::steady_clock::now() - t0)
                                                  .count());
        gr.key = key;
    }
    StepSlot::StepGraph &gr = sl.graphs[gi];
    gr.used = ++sl.submits;
    sl.cur = gi;
    sl.g_blocks = gr.blocks;
    sl.g_tc = gr.tc;
    sl.g_fuse = gr.fuse;
    sl.g_unroll = gr.unroll;
    sl.g_serve = gr.serve;
    sl.g_screen = gr.screen;
    sl.g_agg_full = gr.agg_full;
    sl.step_dense = gr.dense;
    // ---- stage this step (the slot is idle) and commit it to the DB: one
    // record per clause (the device copies it to its bucket row)
    {
        auto fill_part = [&](int k) {
            int64_t *fill = part_row.data() + k * stride;
            uint32_t pos = part_pos[k];
            const int64_t c0 = n * k / cparts, c1 = n * (k + 1) / cparts;
            for (int64_t i = c0; i < c1; i++) {
                const int len = lens[i];
                const Bucket &b = e->buckets[len];
                const int64_t row = b.count + fill[len]++;
                ClauseSeg &c = sl.h_cls.p[i];
                c.dst = b.d + (size_t)row * len;
                c.src = pos;
                c.len = (uint32_t)len;
                if (out_rows) out_rows[i] = (uint32_t)row;
                pos += (uint32_t)len;
            }
        };
        if (cparts > 1)
            pool->run(fill_part);
        else
            fill_part(0);
        for (int len = 1; len <= maxlen; len++) e->buckets[len].count += per[len];
    }
    e->total_clauses += n;
    e->total_lits += total;
    if (n) e->max_lit = std::max<int64_t>(e->max_lit, mx);
    if (n_patches)
        std::memcpy(sl.h_patches.p, desc->patches, (size_t)n_patches * sizeof(gss_var_patch));
    if (n_bups)
        std::memcpy(sl.h_bups.p, desc->base_ups, (size_t)n_bups * sizeof(gss_base_update));
    if (n_loads) std::memcpy(sl.h_loads.p, desc->loads, (size_t)n_loads * sizeof(gss_word_load));
    // descriptors of every non-empty bucket, the step's own inserts included
    int k = 0;
    int64_t t = 0;
    const size_t nb_len = e->buckets.size();
    for (size_t i = 1; i < nb_len; i++) {
        const size_t len = e->tune_schedule == 2 ? i : nb_len - i;
        const Bucket &b = e->buckets[len];
        if (!b.count) continue;
        BucketDesc &d = sl.h_desc.p[k++];
        d.lits = b.d;
        d.len = (uint32_t)len;
        d.count = (uint32_t)b.count;
        d.task_begin = (uint32_t)t;
        d.pad = 0;
        t += (b.count + tc - 1) / tc;
    }
    StepHeader &h = *sl.h_hdr.p;
    std::memset(&h, 0, sizeof h);
    h.nbuckets = k;
    // with no valid slot every clause starts dead: the sweep has no work (a
    // serving step's admission decides that on the device)
    h.ntasks = active_words ? (int32_t)t : 0;
    h.nclauses = (int32_t)n;
    h.nups = n_ups;
    h.npatches = n_patches;
    h.nbups = n_bups;
    h.nloads = n_loads;
    h.region_words[kRegHeader] = sizeof(StepHeader) / 4;
    h.region_words[kRegDesc] = (int64_t)k * (sizeof(BucketDesc) / 4);
    h.region_words[kRegClauses] = n * (int64_t)(sizeof(ClauseSeg) / 4);
    h.region_words[kRegLits] = total;
    h.region_words[kRegUps] = n_ups * (int64_t)(sizeof(gss_var_update) / 4);
    h.region_words[kRegPatches] = n_patches * (int64_t)(sizeof(gss_var_patch) / 4);
    h.region_words[kRegBaseUps] = n_bups * (int64_t)(sizeof(gss_base_update) / 4);
    h.region_words[kRegLoads] = n_loads * (int64_t)(sizeof(gss_word_load) / 4);
    for (int w = 0; w < GSS_MAX_WORDS; w++) {
        h.valid[w] = w < W ? valid[w] : 0;
        h.last_pos[w] = (w < W && last_pos) ? last_pos[w] : -1;
    }
    e->state_version++;
    e->table_agg_at = screen ? e->state_version : 0;
    HIP_TRY(hipEventRecord(sl.t0, e->stream));
    HIP_TRY(hipGraphLaunch(sl.graphs[sl.cur].exec, e->stream));
    HIP_TRY(hipEventRecord(sl.t1, e->stream));
    sl.busy = true;
    sl.settled = false;
    sl.records_lost = false;
    sl.serve = serve;
    // the previous records of this slot stay readable only if their copy
    // was queued (or there were none)
    if (sl.rec_total > 0 && !sl.copy_queued) sl.results = false;
    e->inflight.push_back(S);
    e->next_slot ^= 1;
    return GSS_OK;
}

extern "C" {

int gss_step_wait(gss_engine *e, int64_t *n_hits, float *step_ms) {
    int rc = check_engine(e);
    if (rc) return rc;
    if (e->inflight.empty()) return set_err(GSS_E_STATE, "no step in flight");
    const int S = e->inflight.front();
    StepSlot &sl = e->slots[S];
    HIP_TRY(hipEventSynchronize(sl.t1));
    uint32_t used = sl.h_out.p[0];
    if (used > sl.chunk_cap && !sl.records_lost) {
        // the records overflowed the buffer: re-run the sweep with room
        rc = recover_step(e, 0, used);
        if (rc && !sl.records_lost) {
            e->inflight.erase(e->inflight.begin());
            sl.busy = false;
            return rc;
        }
        HIP_TRY(hipEventSynchronize(sl.t1));
        used = sl.h_out.p[0];
    }
    e->inflight.erase(e->inflight.begin());
    sl.busy = false;
    if (sl.records_lost || used > sl.chunk_cap) {
        // no records to read for this step (nor, until the next wait, any other)
        sl.results = false;
        e->last_step_slot = -1;
        return sl.records_lost
                   ? set_err(GSS_E_NOMEM, "the step's hit buffer overflowed and could not grow: "
                                          "its records are lost (the table did roll forward)")
                   : set_err(GSS_E_NOMEM, "step hit buffer overflow (%u > %u chunks)", used,
                             sl.chunk_cap);
    }
    // the next steps' buffers hold twice the most chunks a step has claimed
    e->step_hit_want = std::max<uint64_t>(e->step_hit_want, 2 * (uint64_t)used);
    int64_t tot = 0;
    for (uint32_t c = 0; c < used; c++) tot += sl.h_out.p[1 + c];
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, sl.t0, sl.t1));
    sl.results = false;
    sl.host_ready = false;
    sl.rec_total = tot;
    sl.res_words = sl.rec_words;
    // the slot's next capture compacts into a buffer with room to spare once
    // the records pass half of the current one (a step whose records do not
    // fit is compacted, or grouped, outside the graph)
    if (2 * (uint64_t)tot > sl.step_dense)
        sl.dense_want = std::max<uint64_t>(sl.dense_want, 4 * (uint64_t)tot);
    e->last_step_slot = S;
    e->last_step_hits = tot;
    e->last_step_chunks = used;
    e->last_step_ms = ms;
    if (tot > 0 && (rc = queue_records_copy(e, sl, tot, used))) return rc;
    sl.results = true;
    if (n_hits) *n_hits = tot;
    if (step_ms) *step_ms = ms;
    return GSS_OK;
}

int gss_step_hits_copy(gss_engine *e, gss_hit *out, int64_t cap, int64_t *n) {
    const gss_hit *p = nullptr;
    int64_t total = 0;
    int rc = gss_step_hits_map(e, &p, &total);
    if (rc) return rc;
    if (n) *n = total;
    const int64_t want = std::min<int64_t>(cap, total);
    if (want <= 0) return GSS_OK;
    if (!out) return set_err(GSS_E_INVAL, "null output");
    std::memcpy(out, p, (size_t)want * sizeof(gss_hit));
    return GSS_OK;
}

int gss_step_hits_map(gss_engine *e, const gss_hit **out, int64_t *n) {
    int rc = check_engine(e);
    if (rc) return rc;
    if (!out) return set_err(GSS_E_INVAL, "null output");
    *out = nullptr;
    if (e->last_step_slot < 0 || !e->slots[e->last_step_slot].results)
        return set_err(GSS_E_STATE, "no waited step with results (the slot was reused)");
    StepSlot &sl = e->slots[e->last_step_slot];
    if (sl.res_words != 3)
        return set_err(GSS_E_STATE, "the last waited step recorded ids or clauses only "
                                    "(gss_step_ids_map, gss_step_clauses_map)");
    const int64_t total = e->last_step_hits;
    if (n) *n = total;
    if (!sl.host_ready && total > 0) HIP_TRY(hipEventSynchronize(sl.ev_copied));
    sl.host_ready = true;
    *out = sl.h_hits().p;
    return GSS_OK;
}

int gss_step_ids_map(gss_engine *e, const gss_hit_id **out, int64_t *n) {
    int rc = check_engine(e);
    if (rc) return rc;
    if (!out) return set_err(GSS_E_INVAL, "null output");
    *out = nullptr;
    if (e->last_step_slot < 0 || !e->slots[e->last_step_slot].results)
        return set_err(GSS_E_STATE, "no waited step with results (the slot was reused)");
    StepSlot &sl = e->slots[e->last_step_slot];
    if (sl.res_words != 1)
        return set_err(GSS_E_STATE, "the last waited step recorded another form");
    const int64_t total = e->last_step_hits;
    if (n) *n = total;
    if (!sl.host_ready && total > 0) HIP_TRY(hipEventSynchronize(sl.ev_copied));
    sl.host_ready = true;
    *out = (const gss_hit_id *)sl.h_hits().p;
    return GSS_OK;
}

int gss_step_clauses_map(gss_engine *e, const gss_hit_clause **out, int64_t *n) {
    int rc = check_engine(e);
    if (rc) return rc;
    if (!out) return set_err(GSS_E_INVAL, "null output");
    *out = nullptr;
    if (e->last_step_slot < 0 || !e->slots[e->last_step_slot].results)
        return set_err(GSS_E_STATE, "no waited step with results (the slot was reused)");
    StepSlot &sl = e->slots[e->last_step_slot];
    if (sl.res_words != 2)
        return set_err(GSS_E_STATE, "the last waited step recorded another form");
    const int64_t total = e->last_step_hits;
    if (n) *n = total;
    if (!sl.host_ready && total > 0) HIP_TRY(hipEventSynchronize(sl.ev_copied));
    sl.host_ready = true;
    *out = (const gss_hit_clause *)sl.h_hits().p;
    return GSS_OK;
}

int gss_db_reserve(gss_engine *e, int64_t n, const int32_t *lens) {
    int rc = check_engine(e);
    if (rc) return rc;
    if ((rc = settle_steps(e))) return rc;
    if (n < 0 || (n > 0 && !lens)) return set_err(GSS_E_INVAL, "bad clause batch");
    if (e->tagged && n > 0) return set_err(GSS_E_STATE, "tagged DB");
    int maxlen = 0;
    for (int64_t i = 0; i < n; i++) {
        if (lens[i] < 1 || lens[i] > GSS_MAX_CLAUSE_LEN)
            return set_err(GSS_E_INVAL, "clause %lld: length %d outside [1, %d]", (long long)i,
                           lens[i], GSS_MAX_CLAUSE_LEN);
        maxlen = std::max(maxlen, (int)lens[i]);
    }
    std::vector<int64_t> per(maxlen + 1, 0);
    for (int64_t i = 0; i < n; i++) per[lens[i]]++;
    if ((int)e->buckets.size() <= maxlen) e->buckets.resize(maxlen + 1);
    for (int len = 1; len <= maxlen; len++)
        if (per[len] && (rc = ensure_bucket_cap(e, len, e->buckets[len].count + per[len])))
            return rc;
    return GSS_OK;
}

}  // extern "C"
