// Device-side data structures and kernel launchers shared by the engine.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gss.h"

namespace gss {

// Hit records are claimed by each wave in chunks of this many records.
constexpr int kHitChunk = 1024;

// One non-empty length bucket, as seen by the sweep kernel.
struct BucketDesc {
    const uint32_t *lits;  // row-major: literal i of row r at lits[r * len + i], stored as
                           // row offsets (literal << 7)
    uint32_t len;
    uint32_t count;        // rows
    uint32_t task_begin;   // first task index of this bucket
    uint32_t pad;
};

struct SweepArgs {
    const BucketDesc *buckets;
    int nbuckets;
    int ntasks;
    int task_clauses;        // clauses per task (multiple of clauses per wave pass)
    const uint64_t *table;   // [2 * vars + 2][words]: row l = slots where literal l is not
                             // false; rows 2*vars (zeros), 2*vars+1 (ones) are sentinels
    const uint64_t *valid;   // [words]
    gss_hit *hits;           // chunk_cap * kHitChunk records
    uint32_t *fills;         // [chunk_cap] records used in each claimed chunk
    uint32_t *alloc;         // number of chunks claimed (may exceed chunk_cap)
    uint32_t chunk_cap;
    int fuse_len;            // clauses of at most this many literals use the fused schedule
                             // (<= 0: none, and the kernel without that path runs)
    uint32_t sentinel;       // stored row offset of the sentinel literal
    const int32_t *dyn;      // if set, {nbuckets, ntasks} are read here (graph replays)
    int rec_words;           // 8-byte words per record: 3 (gss_hit), 1 (gss_hit_id) or 2
                             // (gss_hit_clause: one record per clause, its words as a mask)
    // Serving and screened batched steps (null otherwise): bit b of agg[l]
    // is set only if literal l is non-false in every valid slot of group b --
    // serving: the step's admitted requests r with r % 8 = b; screened steps:
    // word group b, the row all ones over its slots (launch_table_agg);
    // 2 * vars + 2 bytes (padded to 4).  agg_words[b] = the table words of
    // group b (serving: of its requests; screened: with a valid slot).  In a
    // group where two literal occurrences of a clause are so marked no slot
    // can report, so the exact check reads table rows only for the clause's
    // other words.  4 MB at 2M variables: the votes are served from L2.
    const uint8_t *agg;
    const uint64_t *agg_words;  // [8]
    int agg_vote;               // aggregate entries read together per vote: 1, 2, 4 or 8
};

// Staged regions of a batched step, copied host -> device by k_stage_in.
enum StepRegion {
    kRegHeader = 0,
    kRegDesc,
    kRegClauses,
    kRegLits,
    kRegUps,
    kRegPatches,
    kRegBaseUps,
    kRegLoads,
    kStageRegions
};

// Per-step inputs of a batched step (gss_step_*), written by the host into a
// pinned staging slot and copied to the device by k_stage_in.
struct StepHeader {
    int32_t nbuckets;
    int32_t ntasks;
    int32_t nclauses;
    // 1 in the device copy when the step found the engine halted: every
    // count is then 0 and the step does nothing (it is replayed later)
    int32_t halted;
    int64_t nups, npatches, nbups, nloads;
    int64_t region_words[kStageRegions];  // uint32 words of each region to copy
    uint64_t valid[GSS_MAX_WORDS];
    int8_t last_pos[GSS_MAX_WORDS];
};

struct StageArgs {
    const uint32_t *h[kStageRegions];  // pinned host (h[kRegHeader] = the header)
    uint32_t *d[kStageRegions];        // device
    // device word, nonzero while the engine is halted: an earlier step's
    // records overflowed its hit buffer and its sweep is to be re-run before
    // the table moves on
    const uint32_t *halt;
};

struct ScatterSeg {
    uint32_t *dst;
    uint64_t src;    // offset in the staging buffer (uint32 units)
    uint64_t count;  // uint32 elements
};

// One clause staged by a batched step: its literals at src (uint32 units) of
// the staged literals, copied to its bucket row at dst.
struct ClauseSeg {
    uint32_t *dst;
    uint32_t src;
    uint32_t len;
};

// ---- serving: requests admitted on the device -------------------------------

// One request of the serving ring (fine-grained pinned host memory, written
// by the host before it publishes the request).
struct ServeReq {
    uint32_t solver;     // base the request's words are loaded from
    uint32_t k;          // assignments: slots [0, k) of words [0, ceil(k / 64))
    uint64_t patch_off;  // first patch (running index; ring position = modulo capacity)
    uint32_t n_patches;
    uint32_t pad;
    uint64_t seq;        // request sequence number
};

// A request admitted by a step: where it landed and where its patches are.
struct ServeAdmitted {
    uint64_t patch_off;
    uint64_t patch_pfx;  // patches of the step's earlier admitted requests
    uint32_t n_patches;
    uint32_t w0;         // first table word
};

// What a step admitted (pinned host memory, read after the step).
struct ServeReport {
    uint64_t first;  // sequence number of the first admitted request
    uint32_t count;  // requests admitted (in sequence order)
    uint32_t words;  // table words they took
};

// The serving kernels' view of the rings and the device-side cursor.
struct ServeRing {
    const ServeReq *reqs;          // pinned ring, req_cap entries
    const gss_var_patch *patches;  // pinned ring, patch_cap entries (words relative to the request)
    const uint64_t *head;          // pinned: requests published
    uint64_t *tail;                // device: next request to admit
    ServeAdmitted *adm;            // device: [kServeMaxRequests]
    uint32_t *adm_count;           // device
    gss_var_patch *dpatches;       // device: the step's admitted patches, in admission order
                                   // (patch_cap entries)
    uint32_t req_cap;
    uint32_t pad;
    uint64_t patch_cap;
};
constexpr int kServeMaxRequests = 64;  // one admission wave: at most 64 requests per step

// Admits, in sequence order, the published requests whose words fit the
// table (FIFO: the first that does not fit waits for the next step): writes
// the step's word loads, valid masks, load and patch counts (and no tasks
// when nothing was admitted) into the device header, the admitted list, and
// the report.  One wave.
hipError_t launch_serve_admit(const ServeRing &r, StepHeader *dh, gss_word_load *loads,
                              ServeReport *report, int words, hipStream_t s);
// Aggregate entries of a batched step from the table itself (see
// SweepArgs::agg): group b = words [b * words / 8, (b + 1) * words / 8) (one
// word per group below 8 words); bit b of agg[l] set if row l is all ones on
// every slot of group b's words (valid or not: all ones there implies all
// ones on the valid slots); agg_words[b] = the group's words with a valid
// slot in this step.  Reads the table once.
hipError_t launch_table_agg(const uint64_t *table, const StepHeader *dh, int64_t vars, int words,
                            uint8_t *agg, uint64_t *agg_words, hipStream_t s);
// Keeps such entries current without reading the table: for each of the
// step's updates (and patches, if with_patches) the bit of the update's word
// group in both literals of its variable, from the table as it is when the
// kernel runs; agg_words as above if not null.  Run after the updates are
// applied and again after the roll forward.
hipError_t launch_table_agg_update(const uint64_t *table, const StepHeader *dh,
                                   const gss_var_update *ups, const gss_var_patch *patches,
                                   bool with_patches, int words, uint8_t *agg,
                                   uint64_t *agg_words, hipStream_t s);
// Copies the admitted requests' patches from the pinned ring to r.dpatches
// (consecutive lanes read consecutive 4-byte words: one pass over PCIe).
hipError_t launch_serve_stage_patches(const ServeRing &r, const StepHeader *dh, hipStream_t s);
// Builds the aggregate entries of a serving step (see SweepArgs::agg) and
// agg_words from the solver bases its admitted requests were loaded from:
// one pass over the variables, after the admission and the base updates.
hipError_t launch_serve_agg(const ServeRing &r, const StepHeader *dh, const gss_word_load *loads,
                            const uint64_t *bases, int64_t vw, int64_t vars, uint8_t *agg,
                            uint64_t *agg_words, hipStream_t s);
// Applies the admitted requests' patches (staged in r.dpatches) with their
// words shifted to where each request landed; with agg, clears the group bit
// of every literal a patch makes false in a valid slot.
hipError_t launch_serve_patch(const ServeRing &r, const StepHeader *dh, uint64_t *table,
                              int words, uint8_t *agg, hipStream_t s);

// One non-empty bucket of a tagged DB during compaction.
struct CompactRef {
    const uint32_t *lits;  // source rows (stored offsets)
    const uint32_t *tags;
    uint32_t *dst_lits;    // kept rows, in order
    uint32_t *dst_tags;
    int64_t start;         // first row of this bucket in the concatenation
    uint32_t len;
    uint32_t pad;
};

// words is the table row width (power of two, 1..64); unroll selects the
// number of rows in flight per lane for 16 words and wider (0 = default).
hipError_t launch_sweep(const SweepArgs &a, int words, int unroll, int blocks, hipStream_t s);
int sweep_clauses_per_pass(int words);
// Max resident 256-thread blocks per CU for the sweep kernel variant (with or
// without the fused schedule; mode 0: plain, 1: screened, 2: clause records,
// 3: screened with clause records).
int sweep_occupancy(int words, int unroll, bool fused, int mode = 0);

hipError_t launch_reset_u32(uint32_t *p, int n, hipStream_t s);
hipError_t launch_scatter(const ScatterSeg *segs, int nseg, const uint32_t *src, int shift,
                          hipStream_t s);
hipError_t launch_assig_apply(const gss_var_update *ups, int64_t n, uint64_t *table, int words,
                              hipStream_t s);
hipError_t launch_assig_roll(const gss_var_update *ups, int64_t n, uint64_t *table, int words,
                             const int8_t *last_pos, hipStream_t s);
hipError_t launch_fill_u64(uint64_t *p, int64_t n, uint64_t v, hipStream_t s);

// Fixed-shape launches of a batched step (counts read from the device header,
// so a captured graph can be replayed with new data).  While the engine is
// halted k_stage_in writes a header with no work and `halted` set instead of
// staging: every later kernel of the step then has nothing to do (the
// serving kernels and k_table_agg, whose work is not counted by the header,
// test `halted` themselves).
hipError_t launch_stage_in(const StageArgs &a, hipStream_t s);
// Copies each staged clause (n read on the device) to its bucket row,
// shifting literals left by `shift`.
hipError_t launch_scatter_clauses(const ClauseSeg *cls, const int32_t *n, const uint32_t *src,
                                  int shift, hipStream_t s);
hipError_t launch_assig_apply_dyn(const gss_var_update *ups, const int64_t *n, uint64_t *table,
                                  int words, hipStream_t s);
// Rolls the step's updated entries forward -- unless the sweep before it
// claimed more than `cap` hit chunks: then the table stays as the sweep read
// it and *halt is set, so that the sweep can be re-run with room for every
// record.
hipError_t launch_assig_roll_dyn(const gss_var_update *ups, const int64_t *n, uint64_t *table,
                                 int words, const int8_t *last_pos, const uint32_t *alloc,
                                 uint32_t cap, uint32_t *halt, hipStream_t s);
hipError_t launch_stage_out(const uint32_t *alloc, const uint32_t *fills, uint32_t cap,
                            uint32_t *h_out, hipStream_t s);
// Chunks claimed and records of a sweep to pinned host memory: h_out[0] =
// claims, h_out[1] / h_out[2] = low / high word of the record count.
hipError_t launch_count_records(const uint32_t *alloc, const uint32_t *fills, uint32_t cap,
                                uint32_t *h_out, hipStream_t s);
// Dense copy of a sweep's records (chunk order, the records of a chunk in
// place): offs = exclusive prefix of the fills of the claimed chunks (cap + 1
// entries), dense = the records, when they all fit in dense_cap records;
// counts are read on the device.  Records are rec_words 8-byte words (3: a
// gss_hit, 1: a gss_hit_id, 2: a gss_hit_clause); chunk c starts at word
// c * kHitChunk * rec_words.
hipError_t launch_compact_records(const uint32_t *alloc, const uint32_t *fills, uint32_t cap,
                                  uint64_t *offs, const void *hits, void *dense,
                                  uint64_t dense_cap, int rec_words, hipStream_t s);
// Dense copy of a sweep's records grouped by assignment word, ascending:
// counts = 2 * 64 uint32 of device scratch (cleared here), group = 66 uint64:
// group[w] = first record of word w, group[64] = all, group[65] = 1 if they
// fitted in dense_cap records (else nothing moved).
hipError_t launch_group_by_word(const uint32_t *alloc, const uint32_t *fills, uint32_t cap,
                                const void *hits, int rec_words, uint32_t *counts,
                                void *dense, uint64_t dense_cap, uint64_t *group, hipStream_t s);
// Solver base planes: bases[(2 b) * vw + i] / bases[(2 b + 1) * vw + i] hold
// bit (v mod 64) of word i = v / 64 of base b's is_def / is_true.
hipError_t launch_base_apply_dyn(const gss_base_update *u, const int64_t *n, uint64_t *bases,
                                 int64_t vw, hipStream_t s);
// Sets word w of every variable's two rows from base b, for each load.
hipError_t launch_words_load_dyn(const gss_word_load *loads, const int64_t *n,
                                 const uint64_t *bases, int64_t vw, uint64_t *table, int words,
                                 int64_t vars, hipStream_t s);
// Serving steps (loads[k].word == k, words the step does not load invalid in
// it): sets words [0, n) of every variable's two rows from their loads and
// words [n, L) to all ones, L = n rounded up to 16 (at most words), so rows
// are written in whole lines.  words even; vw = (vars + 63) / 64.
hipError_t launch_words_load_lines(const gss_word_load *loads, const int64_t *n,
                                   const uint64_t *bases, int64_t vw, uint64_t *table, int words,
                                   int64_t vars, hipStream_t s);
hipError_t launch_assig_patch_dyn(const gss_var_patch *p, const int64_t *n, uint64_t *table,
                                  int words, hipStream_t s);

// Compaction of a tagged DB (rows of every bucket concatenated, N rows):
// flags[g] = keep[tag of row g] (tags >= ntags are kept), an exclusive scan
// of the flags, the kept count of each bucket, then the moves.
hipError_t launch_compact_flags(const CompactRef *refs, int nb, int64_t n, const uint8_t *keep,
                                int64_t ntags, uint32_t *flags, hipStream_t s);
hipError_t scan_exclusive_u32(const uint32_t *in, uint32_t *out, int64_t n, void *temp,
                              size_t &temp_bytes, hipStream_t s);
hipError_t launch_compact_bounds(const CompactRef *refs, int nb, int64_t n, const uint32_t *flags,
                                 const uint32_t *pos, uint32_t *kept, hipStream_t s);
hipError_t launch_compact_move(const CompactRef *refs, int nb, int64_t n, const uint32_t *flags,
                               const uint32_t *pos, hipStream_t s);

}  // namespace gss
